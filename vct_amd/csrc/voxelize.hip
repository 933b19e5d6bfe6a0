// voxelize.hip -- the dominant-axis voxelizer (GS path) on the binned
// rasterizer.
//
// Per triangle (setup kernel): voxelize.vert (world position, normal matrix),
// voxelize.geom (dominant axis from the summed vertex normals, mvp_axis),
// snap + edge setup.  Per (triangle, pixel) item: 4x MSAA coverage, varyings
// at the pixel centre, the optional per-fragment lighting with 5-tap PCF per
// shadow-casting light, getVoxelPosition's swizzle, and the two packed RGBA8
// imageAtomicMax stores (shaders/voxelize.frag:186-279).  Integer atomicMax on
// the packed word makes the result independent of fragment order.
#pragma clang fp contract(off)

#include "raster_common.h"
#include "texture.h"

namespace vctk {

struct VoxMats {
    float mvp[3][16];
    int axis_override, dim;
};

// attribute block per triangle: 8 float4
//  f[0..8]  ndc xyz of v0,v1,v2        f[9..17]  world positions
//  f[18..26] world normals             f[27] albedo bits  f[28] axis | zclip<<4
//  f[29] rho^2 of the diffuse map's footprint (constant over the triangle: the
//  dominant-axis views are orthographic)  f[30] the map's index, -1 for none
__global__ __launch_bounds__(256) void k_vox_setup(DevScene sc, VoxMats vm, int4 *rec, float4 *attr,
                                                   uint32_t *counts) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= sc.ntri) return;
    const int o = (int)sc.obj[t];
    const float *M = sc.model + (size_t)o * 16;
    const float *Nm = sc.nmat + (size_t)o * 9;
    float W[3][3], Nw[3][3];
    for (int v = 0; v < 3; ++v) {
        const float *p = sc.pos + (size_t)t * 9 + v * 3;
        float p4[4] = {p[0], p[1], p[2], 1.0f}, w4[4];
        mat4_vec4(M, p4, w4);
        W[v][0] = w4[0];
        W[v][1] = w4[1];
        W[v][2] = w4[2];
        mat3_vec3(Nm, sc.nrm + (size_t)t * 9 + v * 3, Nw[v]);
    }
    float fs[3], fn[3];
    for (int k = 0; k < 3; ++k) fs[k] = (Nw[0][k] + Nw[1][k]) + Nw[2][k];
    nrm3(fs, fn);
    const float ax = __builtin_fabsf(fn[0]), ay = __builtin_fabsf(fn[1]), az = __builtin_fabsf(fn[2]);
    int axis;
    if (ax > ay && ax > az) axis = 0;
    else if (ay > ax && ay > az) axis = 1;
    else axis = 2;
    if (vm.axis_override >= 0 && vm.axis_override <= 2) axis = vm.axis_override;
    const float *mvp = vm.mvp[axis];
    const float half = (float)vm.dim * 0.5f;
    float ndc[3][3], win[3][2];
    int zclip = 0;
    for (int v = 0; v < 3; ++v) {
        float w4[4] = {W[v][0], W[v][1], W[v][2], 1.0f}, c[4];
        mat4_vec4(mvp, w4, c);
        for (int k = 0; k < 3; ++k) ndc[v][k] = c[k] / c[3];
        if (ndc[v][2] < -1.0f || ndc[v][2] > 1.0f) zclip = 1;
        win[v][0] = ndc[v][0] * half + half;
        win[v][1] = ndc[v][1] * half + half;
    }
    int4 r[4];
    const uint64_t cnt = tri_setup(win, vm.dim, vm.dim, 0, t, r);
    if (cnt) {
        for (int i = 0; i < 4; ++i) rec[(size_t)t * 4 + i] = r[i];
        float f[32];
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) {
                f[v * 3 + k] = ndc[v][k];
                f[9 + v * 3 + k] = W[v][k];
                f[18 + v * 3 + k] = Nw[v][k];
            }
        f[27] = __uint_as_float(sc.albedo[t]);
        f[28] = __uint_as_float((uint32_t)axis | ((uint32_t)zclip << 4));
        const int tid = sc.texid ? sc.texid[t] : -1;
        float rho2 = 0.0f;
        if (tid >= 0 && tid < sc.tex.count) {
            const int4 inf = sc.tex.info[tid];
            rho2 = tex_rho2_affine(win, sc.uv + (size_t)t * 6, (float)inf.y, (float)inf.z);
        }
        f[29] = rho2;
        f[30] = __int_as_float(tid);
        f[31] = 0.0f;
        float4 *a = attr + (size_t)t * 8;
        for (int i = 0; i < 8; ++i) a[i] = make_float4(f[i * 4], f[i * 4 + 1], f[i * 4 + 2], f[i * 4 + 3]);
    }
    counts[t] = (uint32_t)cnt;
}

// calcShadowFactor (shaders/voxelize.frag:160-184)
// Coverage of one (triangle, pixel) item: 4x MSAA with the per-sample
// near/far clip; on coverage, the pixel-centre barycentrics in l.
__device__ __forceinline__ bool vox_cover(const float4 *attr, int msaa, int slot, int px, int py, const Edges &e,
                                          float l[3]) {
    const float4 *a = attr + (size_t)slot * 8;
        const uint32_t flags = __float_as_uint(a[7].x);
        const bool zclip = (flags >> 4) & 1u;
        // ndc z of the three vertices: f[2], f[5], f[8]
        const float nz0 = a[0].z, nz1 = a[1].y, nz2 = a[2].x;
        bool cov = false;
        const int ns = msaa ? 4 : 1;
        for (int k = 0; k < ns && !cov; ++k) {
            const int ox = msaa ? msaa_x(k) : 128, oy = msaa ? msaa_y(k) : 128;
            if (!edge_inside(e, px, py, ox, oy)) continue;
            if (zclip) {
                float ls[3];
                edge_bary(e, px, py, ox, oy, ls);
                const float z = interp3(ls, nz0, nz1, nz2);
                if (!(z >= -1.0f && z <= 1.0f)) continue;
            }
            cov = true;
        }
        if (cov) edge_bary(e, px, py, 128, 128, l);
        return cov;
}

// getVoxelPosition's swizzle (voxelize.frag:79-101): NDC of the fragment to
// unit texture coordinates of the grid.
__device__ __forceinline__ void vox_unit(const float pos[3], int axis, float un[3]) {
    float u[3];
    for (int k = 0; k < 3; ++k) u[k] = (pos[k] + 1.0f) * 0.5f;
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
}

// imageAtomicRGBA8Avg, the running average the reference's RGBA8 build uses
// when voxelizeAtomicMax is off (voxelize.frag:111-139, the live `#if 1`
// branch): the word holds the average colour in its rgb bytes (0..255) and
// the fragment count in its alpha byte; each fragment adds (255 c, 1) by
// compare-and-swap, rgb = (rgb * count + 255 c) / (count + 1) truncated to
// bytes.  The result depends on the order in which fragments land (the count
// does not).  A lane whose swap fails retries with the value it found, so
// every attempt by some lane succeeds and the loop ends.
__device__ __forceinline__ uint32_t rgba8_from_floats(float x, float y, float z, float w) {
    return (((uint32_t)w & 0xffu) << 24) | (((uint32_t)z & 0xffu) << 16) | (((uint32_t)y & 0xffu) << 8) |
           ((uint32_t)x & 0xffu);
}
__device__ __forceinline__ void rgba8_avg(uint32_t *p, float r, float g, float b) {
    const float v0 = r * 255.0f, v1 = g * 255.0f, v2 = b * 255.0f, v3 = 1.0f;
    uint32_t nv = rgba8_from_floats(v0, v1, v2, v3);
    uint32_t prev = 0u, cur;
    while ((cur = atomicCAS(p, prev, nv)) != prev) {
        prev = cur;
        float c0 = (float)(cur & 0xffu), c1 = (float)((cur >> 8) & 0xffu), c2 = (float)((cur >> 16) & 0xffu),
              c3 = (float)(cur >> 24);
        c0 = c0 * c3;
        c1 = c1 * c3;
        c2 = c2 * c3;
        float s0 = c0 + v0, s1 = c1 + v1, s2 = c2 + v2;
        const float s3 = c3 + v3;
        const double r3 = 1.0 / (double)s3;  // the three IEEE quotients by s3 (quot_f32)
        s0 = quot_f32(s0, r3);
        s1 = quot_f32(s1, r3);
        s2 = quot_f32(s2, r3);
        nv = rgba8_from_floats(s0, s1, s2, s3);
    }
}

// The sum-accumulate mode (the reference's fp16 imageAtomicAdd of (c, 1),
// voxelize.frag:257-260, normalized by normalizeVoxels.comp): here the sums
// are 32-bit fixed point, round(65535 clamp(c, 0, 1)) per channel plus the
// count, so they are exact and independent of fragment order.  Eight words
// per voxel: colour rgb + count, normal rgb + count (voxel index v at
// sums[8 v ..]).
__device__ __forceinline__ uint32_t fix16(float c) { return (uint32_t)__builtin_rintf(clampf(c, 0.0f, 1.0f) * 65535.0f); }

// The scene's lights in LDS, for the lit voxelizer (VoxFrag::begin).
__shared__ DevLight vox_lights[kMaxLights];

struct VoxFrag {
    const float4 *attr;
    const DevLight *lights;
    int nlights, lighting, msaa, dim;
    int accum;        // VCT_ACCUM_* (voxelize params atomic_max)
    uint32_t *sums;   // VCT_ACCUM_SUM: 8 words per voxel
    uint32_t *albedo, *normal;
    uint32_t *partials;
    const float4 *warp;  // warpTexture: the warp texture, or null
    const float *uv;     // the scene's texture coordinates and diffuse maps
    DevTex tex;
    uint32_t frags;

    // The lights are read field by field inside the shading loop, behind
    // branches on their type and flags: from LDS (vox_lights) each such read
    // is a short round trip instead of one to the cache hierarchy.
    __device__ void begin() {
        const int nw = (lighting ? nlights : 0) * (int)(sizeof(DevLight) / 4);
        for (int i = threadIdx.x; i < nw; i += blockDim.x)
            ((uint32_t *)vox_lights)[i] = ((const uint32_t *)lights)[i];
        __syncthreads();
    }
    __device__ bool cover(int slot, int px, int py, const Edges &e, float l[3]) const {
        return vox_cover(attr, msaa, slot, px, py, e, l);
    }

    // One covered fragment: varyings, lighting and the two atomicMax stores.
    __device__ void shade(int slot, const float l[3]) {
        const float4 *a = attr + (size_t)slot * 8;
        const uint32_t flags = __float_as_uint(a[7].x);
        const int axis = (int)(flags & 3u);
        const float4 a0 = a[0], a1 = a[1], a2 = a[2];
        const float4 a3 = a[3], a4 = a[4], a5 = a[5], a6 = a[6];
        float f[28];
        f[0] = a0.x; f[1] = a0.y; f[2] = a0.z; f[3] = a0.w;
        f[4] = a1.x; f[5] = a1.y; f[6] = a1.z; f[7] = a1.w;
        f[8] = a2.x; f[9] = a2.y; f[10] = a2.z; f[11] = a2.w;
        f[12] = a3.x; f[13] = a3.y; f[14] = a3.z; f[15] = a3.w;
        f[16] = a4.x; f[17] = a4.y; f[18] = a4.z; f[19] = a4.w;
        f[20] = a5.x; f[21] = a5.y; f[22] = a5.z; f[23] = a5.w;
        f[24] = a6.x; f[25] = a6.y; f[26] = a6.z; f[27] = a6.w;
        float pos[3], wp[3], nr[3];
        for (int k = 0; k < 3; ++k) {
            pos[k] = interp3(l, f[k], f[3 + k], f[6 + k]);
            wp[k] = interp3(l, f[9 + k], f[12 + k], f[15 + k]);
            nr[k] = interp3(l, f[18 + k], f[21 + k], f[24 + k]);
        }
        frags += 1;
        const uint32_t alb = __float_as_uint(f[27]);
        float color[3] = {unorm8(alb, 0), unorm8(alb, 1), unorm8(alb, 2)};
        const float4 a7 = a[7];
        const int tid = __float_as_int(a7.z);
        if (tid >= 0) {  // texture(diffuseMap, texcoord).rgb (voxelize.frag:197)
            const float *st = uv + (size_t)slot * 6;
            const float ts = interp3(l, st[0], st[2], st[4]), tt = interp3(l, st[1], st[3], st[5]);
            const float4 c = tex_sample(tex, tid, ts, tt, a7.y);
            color[0] = c.x;
            color[1] = c.y;
            color[2] = c.z;
        }
        float nn[3];
        nrm3(nr, nn);
        const float ne0 = (nn[0] + 1.0f) * 0.5f, ne1 = (nn[1] + 1.0f) * 0.5f, ne2 = (nn[2] + 1.0f) * 0.5f;
        if (lighting) {
            // Two lights per pass: each light's term first, then the shadow
            // taps of both lights fetched before either is compared, so their
            // memory latency overlaps -- and only for a light whose term is
            // non-zero in some lane of the wavefront: a zero term (facing
            // away, beyond a point light's range) stays exactly zero whatever
            // the shadow factor (finite, in [0, 1]).  Each light's term and
            // the sum run in light order, as one light at a time would.
            float fin[3] = {0.0f, 0.0f, 0.0f};
            for (int i = 0; i < nlights; i += 2) {
                ShadowTaps st[2];
                bool fetch[2];
                float lit[2][3];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    fetch[h] = false;
                    lit[h][0] = lit[h][1] = lit[h][2] = 0.0f;
                    if (i + h >= nlights) continue;
                    const DevLight &L = vox_lights[i + h];
                    if (!L.enabled) continue;
                    if (L.type == 0u) {
                        float d[3] = {L.position[0] - wp[0], L.position[1] - wp[1], L.position[2] - wp[2]};
                        float dn[3];
                        nrm3(d, dn);
                        const float ndl = __builtin_fmaxf(0.0f, dot3(nn, dn));
                        for (int k = 0; k < 3; ++k) lit[h][k] = ((color[k] * L.intensity) * L.color[k]) * ndl;
                        const float dist = __builtin_sqrtf(dot3(d, d));
                        const float e0 = 0.75f * L.range, e1 = L.range;
                        const float tt = clampf((dist - e0) / (e1 - e0), 0.0f, 1.0f);
                        const float ss = (tt * tt) * (3.0f - 2.0f * tt);
                        const float att = 1.0f - ss;
                        for (int k = 0; k < 3; ++k) lit[h][k] = lit[h][k] * att;
                    } else {
                        float nd[3] = {-L.direction[0], -L.direction[1], -L.direction[2]}, ndn[3];
                        nrm3(nd, ndn);
                        const float ndl = __builtin_fmaxf(0.0f, dot3(nn, ndn));
                        for (int k = 0; k < 3; ++k) lit[h][k] = ((color[k] * L.color[k]) * L.intensity) * ndl;
                    }
                    const bool nz = lit[h][0] != 0.0f || lit[h][1] != 0.0f || lit[h][2] != 0.0f;
                    fetch[h] = L.shadow_caster && L.shadowquad && __ballot(nz) != 0;  // wave-uniform
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (!fetch[h]) continue;
                    const DevLight &L = vox_lights[i + h];
                    float w4[4] = {wp[0], wp[1], wp[2], 1.0f}, lsp[4];
                    mat4_vec4(L.ls, w4, lsp);
                    shadow_taps(lsp, L.shadowquad, (int)L.shadow_size, st[h]);
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (i + h >= nlights) continue;
                    const DevLight &L = vox_lights[i + h];
                    if (!L.enabled) continue;
                    if (fetch[h]) {
                        const float sf = 1.0f - shadow_count(st[h], (int)L.shadow_size);
                        for (int k = 0; k < 3; ++k) lit[h][k] = lit[h][k] * sf;
                    }
                    for (int k = 0; k < 3; ++k) fin[k] = fin[k] + lit[h][k];
                }
            }
            for (int k = 0; k < 3; ++k) color[k] = clampf(fin[k], 0.0f, 1.0f);
        }
        float un[3];
        vox_unit(pos, axis, un);
        if (warp) warp_sample(warp, un, un);  // voxelize.frag:103-105
        int idx[3];
        for (int k = 0; k < 3; ++k) idx[k] = (int)((float)dim * un[k]);
        if (idx[0] < 0 || idx[1] < 0 || idx[2] < 0 || idx[0] >= dim || idx[1] >= dim || idx[2] >= dim) return;
        const uint32_t vi = grid_index(idx[0], idx[1], idx[2], dim >> 2);
        if (accum == VCT_ACCUM_MAX) {
            atomicMax(&albedo[vi], pack_unorm4x8(color[0], color[1], color[2], 1.0f));
            atomicMax(&normal[vi], pack_unorm4x8(ne0, ne1, ne2, 1.0f));
        } else if (accum == VCT_ACCUM_AVG) {
            rgba8_avg(&albedo[vi], color[0], color[1], color[2]);
            rgba8_avg(&normal[vi], ne0, ne1, ne2);
        } else {
            uint32_t *w = sums + (size_t)vi * 8;
            atomicAdd(&w[0], fix16(color[0]));
            atomicAdd(&w[1], fix16(color[1]));
            atomicAdd(&w[2], fix16(color[2]));
            atomicAdd(&w[3], 1u);
            atomicAdd(&w[4], fix16(ne0));
            atomicAdd(&w[5], fix16(ne1));
            atomicAdd(&w[6], fix16(ne2));
            atomicAdd(&w[7], 1u);
        }
    }
    __device__ void finish() { block_partial(frags, 0u, 0u, 0u, partials); }
};

// The occupancy mode of the voxelizer (voxelize.frag:187-193): a covered
// fragment ORs 1 into the 32^3 cell of its unwarped position.
struct OccFrag {
    const float4 *attr;
    int msaa;
    uint32_t *occ;

    __device__ void begin() {}
    __device__ bool cover(int slot, int px, int py, const Edges &e, float l[3]) const {
        return vox_cover(attr, msaa, slot, px, py, e, l);
    }
    __device__ void shade(int slot, const float l[3]) {
        const float4 *a = attr + (size_t)slot * 8;
        const int axis = (int)(__float_as_uint(a[7].x) & 3u);
        const float4 a0 = a[0], a1 = a[1], a2 = a[2];
        const float pos[3] = {interp3(l, a0.x, a0.w, a1.z), interp3(l, a0.y, a1.x, a1.w),
                              interp3(l, a0.z, a1.y, a2.x)};
        float un[3];
        vox_unit(pos, axis, un);
        int idx[3];
        for (int k = 0; k < 3; ++k) idx[k] = (int)((float)kWarpDim * un[k]);
        if (idx[0] < 0 || idx[1] < 0 || idx[2] < 0 || idx[0] >= kWarpDim || idx[1] >= kWarpDim ||
            idx[2] >= kWarpDim)
            return;
        atomicOr(&occ[(idx[2] * kWarpDim + idx[1]) * kWarpDim + idx[0]], 1u);
    }
    __device__ void finish() {}
};

hipError_t launch_warp_occupancy(const DevScene &sc, const vct_voxelize_params &p, uint32_t *occ, RasterScratch &rs,
                                 hipStream_t s) {
    hipError_t e = launch_fill_u32(occ, 0u, (size_t)kWarpDim * kWarpDim * kWarpDim, s);  // Application.cpp:257
    if (e != hipSuccess || sc.ntri == 0) return e;
    VoxMats vm;
    for (int i = 0; i < 16; ++i) {
        vm.mvp[0][i] = p.mvp_x[i];
        vm.mvp[1][i] = p.mvp_y[i];
        vm.mvp[2][i] = p.mvp_z[i];
    }
    vm.axis_override = p.axis_override;
    vm.dim = kWarpDim;  // glViewport(0, 0, voxelOccupancyDim, voxelOccupancyDim) (:239)
    hipLaunchKernelGGL(k_vox_setup, dim3((sc.ntri + 255) / 256), dim3(256), 0, s, sc, vm, rs.rec, rs.attr, rs.counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = raster_scan(rs, (size_t)sc.ntri, false, s)) != hipSuccess) return e;
    OccFrag f;
    f.attr = rs.attr;
    f.msaa = p.msaa;
    f.occ = occ;
    hipLaunchKernelGGL(k_raster_queued<OccFrag>, dim3(raster_grid<OccFrag, true>()), dim3(256), 0, s, rs.rec, rs.offs,
                       rs.chunk,
                       rs.chunk_cap, sc.ntri, f);
    return hipGetLastError();
}

// glClearTexImage of both voxel grids (Application.cpp:627-628, 686-687): 16
// bytes of each grid per thread.  zero (or null): the zero record, whose bit 2
// marks the bricks that hold 0 in both grids already; they are not rewritten.
__global__ __launch_bounds__(256) void k_clear_grids(uint32_t *albedo, uint32_t *normal, size_t nchunks,
                                                     const uint8_t *__restrict__ zero) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nchunks; i += (size_t)gridDim.x * blockDim.x) {
        if (zero != nullptr && (zero[i >> 4] & 4u) != 0) continue;  // 16 chunks of 16 bytes per 4^3 brick
        ((uint4 *)albedo)[i] = make_uint4(0u, 0u, 0u, 0u);
        ((uint4 *)normal)[i] = make_uint4(0u, 0u, 0u, 0u);
    }
}

hipError_t launch_clear_grids(uint32_t *albedo, uint32_t *normal, int dim, const uint8_t *zero, hipStream_t s) {
    const size_t nchunks = (size_t)dim * dim * dim / 4;
    const size_t want = (nchunks + 255) / 256;
    const unsigned blocks = (unsigned)(want < 8192 ? want : 8192);
    hipLaunchKernelGGL(k_clear_grids, dim3(blocks), dim3(256), 0, s, albedo, normal, nchunks, zero);
    return hipGetLastError();
}

hipError_t launch_voxelize(const DevScene &sc, const DevLight *dlights, int nlights, const vct_voxelize_params &p,
                           int dim, uint32_t *albedo, uint32_t *normal, hipEvent_t grids_ready, uint32_t *sums,
                           vct_stats *stats, uint32_t *partials, const float4 *warp, RasterScratch &rs,
                           hipStream_t s) {
    const size_t nv = (size_t)dim * dim * dim;
    hipError_t e = hipSuccess;
    if (p.atomic_max == VCT_ACCUM_SUM && (e = launch_fill_u32(sums, 0u, nv * 8, s)) != hipSuccess) return e;
    if (sc.ntri == 0) {
        if ((e = hipStreamWaitEvent(s, grids_ready, 0)) != hipSuccess) return e;
        return launch_fill_u32(&stats->total_voxel_fragments, 0u, 1, s);
    }
    VoxMats vm;
    for (int i = 0; i < 16; ++i) {
        vm.mvp[0][i] = p.mvp_x[i];
        vm.mvp[1][i] = p.mvp_y[i];
        vm.mvp[2][i] = p.mvp_z[i];
    }
    vm.axis_override = p.axis_override;
    vm.dim = dim;
    hipLaunchKernelGGL(k_vox_setup, dim3((sc.ntri + 255) / 256), dim3(256), 0, s, sc, vm, rs.rec, rs.attr, rs.counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = raster_scan(rs, (size_t)sc.ntri, false, s)) != hipSuccess) return e;
    VoxFrag f;
    f.attr = rs.attr;
    f.lights = dlights;
    f.nlights = nlights;
    f.lighting = p.lighting;
    f.msaa = p.msaa;
    f.dim = dim;
    f.accum = p.atomic_max;
    f.sums = sums;
    f.albedo = albedo;
    f.normal = normal;
    f.partials = partials;
    f.warp = warp;
    f.uv = sc.uv;
    f.tex = sc.tex;
    f.frags = 0;
    static_assert(kRasterBlocks <= kStatPartials, "one statistics partial per rasterizer workgroup");
    const int nblk = raster_grid<VoxFrag, true>();
    if ((e = hipStreamWaitEvent(s, grids_ready, 0)) != hipSuccess) return e;  // the grids are clear
    hipLaunchKernelGGL(k_raster_queued<VoxFrag>, dim3(nblk), dim3(256), 0, s, rs.rec, rs.offs, rs.chunk,
                       rs.chunk_cap, sc.ntri, f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // totalVoxelFragments (voxelize.frag:195)
    return launch_reduce_stats(partials, nblk, &stats->total_voxel_fragments, nullptr, nullptr, s);
}

// ---------------------------------------------------------------- tessellation voxelizer
// The voxelizer the reference runs when voxelizeTesselation is set
// (Application.cpp:585-654): simpleTesselated.vert -> testTesselation.tesc ->
// testTesselation.tese in point mode with the rasterizer discarded.  Every
// patch (triangle) becomes a set of tessellation points and each point stores
// the triangle's albedo and its interpolated normal in the voxel under it.
//  - .tesc:40-49: a patch whose three vertices fall in one voxel,
//    ivec3(dim * linear position), gets levels 0 and is discarded;
//  - .tesc:51-75: inner level max(1, max altitude / voxelSize.x) (Heron's
//    area), outer level of an edge max(1, length / length(normalize(edge) *
//    voxelSize)), assigned OL0 = b - a, OL1 = c - b, OL2 = c - a as written;
//  - point-mode equal_spacing tessellation of the triangle domain (GL 4.5
//    core, 11.2.2.1): levels clamped to [1, 64] and rounded up, an inner level
//    1 with an outer level above 1 taken as 2, all levels 1 giving just the
//    corners; else the corners, OL_e - 1 points inside outer edge e (u = 0,
//    v = 0, w = 0 for e = 0, 1, 2), and inner triangles k = 1 .. n/2 with
//    corners (1 - 4k/3n, 2k/3n, 2k/3n) and rotations (the perpendiculars
//    through the k-th subdivision points of the edges), each edge cut into
//    n - 2k segments (0: the centre);
//  - .tese:121-140: position and normal interpolated with (u, v, w), voxel
//    ivec3(dim * linear position), colour the triangle's albedo (standing for
//    its diffuse map), normal normalize(n) * 0.5 + 0.5, atomicMax or the RGBA8
//    running average; no lighting, no fragment count.
// A setup pass counts every patch's points (attribute block: world positions
// and normals, the levels in the setup record); the scan spreads the points
// over the threads, one point each.
struct TessArgs {
    int dim;
    float center[3], vmin[3], vmax[3];
};

__device__ __forceinline__ int tess_count(const int lv[4]) {
    const int n = lv[0];
    if (n == 1 && lv[1] == 1 && lv[2] == 1 && lv[3] == 1) return 3;
    int cnt = 3 + (lv[1] - 1) + (lv[2] - 1) + (lv[3] - 1);
    for (int k = 1; 2 * k <= n; ++k) {
        const int m = n - 2 * k;
        cnt += m == 0 ? 1 : 3 * m;
    }
    return cnt;
}

__global__ __launch_bounds__(256) void k_tess_setup(DevScene sc, TessArgs ta, int4 *rec, float4 *attr,
                                                    uint32_t *counts) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= sc.ntri) return;
    const int o = (int)sc.obj[t];
    const float *M = sc.model + (size_t)o * 16;
    const float *Nm = sc.nmat + (size_t)o * 9;
    float P[3][3], N[3][3];
    for (int v = 0; v < 3; ++v) {
        const float *p = sc.pos + (size_t)t * 9 + v * 3;
        float p4[4] = {p[0], p[1], p[2], 1.0f}, w4[4];
        mat4_vec4(M, p4, w4);
        P[v][0] = w4[0];
        P[v][1] = w4[1];
        P[v][2] = w4[2];
        mat3_vec3(Nm, sc.nrm + (size_t)t * 9 + v * 3, N[v]);
    }
    float4 *a = attr + (size_t)t * 8;  // P0 P1 P2 N0 N1 N2 as 18 floats
    a[0] = make_float4(P[0][0], P[0][1], P[0][2], P[1][0]);
    a[1] = make_float4(P[1][1], P[1][2], P[2][0], P[2][1]);
    a[2] = make_float4(P[2][2], N[0][0], N[0][1], N[0][2]);
    a[3] = make_float4(N[1][0], N[1][1], N[1][2], N[2][0]);
    a[4] = make_float4(N[2][1], N[2][2], 0.0f, 0.0f);
    int vox[3][3];
    for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) {
            const float tc = ((P[v][k] - ta.center[k]) - ta.vmin[k]) / (ta.vmax[k] - ta.vmin[k]);
            vox[v][k] = (int)((float)ta.dim * tc);
        }
    bool same = true;
    for (int k = 0; k < 3; ++k) same = same && vox[0][k] == vox[1][k] && vox[0][k] == vox[2][k];
    if (same) {
        counts[t] = 0;
        return;
    }
    float E[3][3];  // A = c - b, B = c - a, C = b - a
    for (int k = 0; k < 3; ++k) {
        E[0][k] = P[2][k] - P[1][k];
        E[1][k] = P[2][k] - P[0][k];
        E[2][k] = P[1][k] - P[0][k];
    }
    float len[3];
    for (int i = 0; i < 3; ++i) len[i] = __builtin_sqrtf(dot3(E[i], E[i]));
    const float sp = ((len[0] + len[1]) + len[2]) * 0.5f;
    const float area = __builtin_sqrtf(((sp * (sp - len[0])) * (sp - len[1])) * (sp - len[2]));
    float maxalt = 0.0f;
    for (int i = 0; i < 3; ++i) {
        const float alt = (2.0f * area) / len[i];
        maxalt = i == 0 ? alt : __builtin_fmaxf(maxalt, alt);
    }
    float vs[3];
    for (int k = 0; k < 3; ++k) vs[k] = (ta.vmax[k] - ta.vmin[k]) / (float)ta.dim;
    float outer[3];
    for (int i = 0; i < 3; ++i) {
        float n[3];
        nrm3(E[i], n);
        const float d[3] = {n[0] * vs[0], n[1] * vs[1], n[2] * vs[2]};
        const float div = __builtin_fabsf(__builtin_sqrtf(dot3(d, d)));
        outer[i] = __builtin_fmaxf(1.0f, len[i] / div);
    }
    const float lev[4] = {__builtin_fmaxf(1.0f, maxalt / vs[0]), outer[2], outer[0], outer[1]};
    int lv[4];
    for (int i = 0; i < 4; ++i) {
        const float x = lev[i] > 64.0f ? 64.0f : lev[i];
        lv[i] = (int)__builtin_ceilf(x);
    }
    if (lv[0] == 1 && (lv[1] > 1 || lv[2] > 1 || lv[3] > 1)) lv[0] = 2;
    rec[(size_t)t * 4] = make_int4(lv[0], lv[1], lv[2], lv[3]);
    counts[t] = (uint32_t)tess_count(lv);
}

// barycentric (u, v, w) of point j of a patch with levels lv (the order of
// the oracle's or_tess_point: corners, outer edges 0..2, inner rings outward in)
__device__ __forceinline__ void tess_point(const int lv[4], int j, float w[3]) {
    if (j < 3) {
        w[0] = j == 0 ? 1.0f : 0.0f;
        w[1] = j == 1 ? 1.0f : 0.0f;
        w[2] = j == 2 ? 1.0f : 0.0f;
        return;
    }
    j -= 3;
    for (int e = 0; e < 3; ++e) {
        const int m = lv[1 + e];
        if (j < m - 1) {
            const float f = (float)(j + 1) / (float)m;
            const float g = 1.0f - f;
            if (e == 0) {
                w[0] = 0.0f;
                w[1] = g;
                w[2] = f;
            } else if (e == 1) {
                w[0] = f;
                w[1] = 0.0f;
                w[2] = g;
            } else {
                w[0] = g;
                w[1] = f;
                w[2] = 0.0f;
            }
            return;
        }
        j -= m - 1;
    }
    const int n = lv[0];
    for (int k = 1; 2 * k <= n; ++k) {
        const int m = n - 2 * k;
        const float cc = (2.0f * (float)k) / (3.0f * (float)n);
        const float oo = 1.0f - 2.0f * cc;
        const int cnt = m == 0 ? 1 : 3 * m;
        if (j < cnt) {
            if (m == 0) {
                w[0] = oo;
                w[1] = cc;
                w[2] = cc;
                return;
            }
            const int e = j / m, s = j % m;
            const float f = (float)s / (float)m;
            // corners R0 = (o, c, c), R1 = (c, o, c), R2 = (c, c, o); edge e: R_e -> R_(e+1)%3
            float ra[3] = {cc, cc, cc}, rb[3] = {cc, cc, cc};
            ra[e] = oo;
            rb[(e + 1) % 3] = oo;
            for (int q = 0; q < 3; ++q) w[q] = ra[q] + f * (rb[q] - ra[q]);
            return;
        }
        j -= cnt;
    }
    w[0] = w[1] = w[2] = 0.0f;
}

__global__ __launch_bounds__(256) void k_tess_points(DevScene sc, TessArgs ta, const int4 *__restrict__ rec,
                                                     const float4 *__restrict__ attr, const uint64_t *__restrict__ offs,
                                                     const uint32_t *__restrict__ chunk, size_t chunk_cap, int accum,
                                                     uint32_t *albedo, uint32_t *normal) {
    const int nslots = sc.ntri;
    const uint64_t total = offs[nslots];
    const uint64_t nchunks = (total + 63) >> 6;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t c = k >> 6;
        int lo = c < chunk_cap ? (int)chunk[c] : 0;
        int hi = (c + 1 < nchunks && c + 1 < chunk_cap) ? (int)chunk[c + 1] : nslots - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (offs[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        const int t = lo;
        const int j = (int)(k - offs[t]);
        const int4 r = rec[(size_t)t * 4];
        const int lv[4] = {r.x, r.y, r.z, r.w};
        float w[3];
        tess_point(lv, j, w);
        const float4 *a = attr + (size_t)t * 8;
        const float4 a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4];
        const float P[3][3] = {{a0.x, a0.y, a0.z}, {a0.w, a1.x, a1.y}, {a1.z, a1.w, a2.x}};
        const float N[3][3] = {{a2.y, a2.z, a2.w}, {a3.x, a3.y, a3.z}, {a3.w, a4.x, a4.y}};
        float p[3], nr[3];
        for (int q = 0; q < 3; ++q) {
            p[q] = (w[0] * P[0][q] + w[1] * P[1][q]) + w[2] * P[2][q];
            nr[q] = (w[0] * N[0][q] + w[1] * N[1][q]) + w[2] * N[2][q];
        }
        int idx[3];
        for (int q = 0; q < 3; ++q)
            idx[q] = (int)((float)ta.dim * (((p[q] - ta.center[q]) - ta.vmin[q]) / (ta.vmax[q] - ta.vmin[q])));
        if (idx[0] < 0 || idx[1] < 0 || idx[2] < 0 || idx[0] >= ta.dim || idx[1] >= ta.dim || idx[2] >= ta.dim)
            continue;
        float nn[3];
        nrm3(nr, nn);
        const float ne0 = nn[0] * 0.5f + 0.5f, ne1 = nn[1] * 0.5f + 0.5f, ne2 = nn[2] * 0.5f + 0.5f;
        const uint32_t alb = sc.albedo[t];
        float4 col = make_float4(unorm8(alb, 0), unorm8(alb, 1), unorm8(alb, 2), unorm8(alb, 3));
        const int tid = sc.texid ? sc.texid[t] : -1;
        if (tid >= 0) {  // texture(diffuseMap, texcoord) in the evaluation shader: zero footprint
            const float *st = sc.uv + (size_t)t * 6;
            const float ts = (w[0] * st[0] + w[1] * st[2]) + w[2] * st[4];
            const float tt = (w[0] * st[1] + w[1] * st[3]) + w[2] * st[5];
            col = tex_sample(sc.tex, tid, ts, tt, 0.0f);
        }
        const uint32_t vi = grid_index(idx[0], idx[1], idx[2], ta.dim >> 2);
        if (accum == VCT_ACCUM_MAX) {
            atomicMax(&albedo[vi], pack_unorm4x8(col.x, col.y, col.z, col.w));
            atomicMax(&normal[vi], pack_unorm4x8(ne0, ne1, ne2, 1.0f));
        } else {
            rgba8_avg(&albedo[vi], col.x, col.y, col.z);
            rgba8_avg(&normal[vi], ne0, ne1, ne2);
        }
    }
}

hipError_t launch_voxelize_tess(const DevScene &sc, const float center[3], const float vmin[3], const float vmax[3],
                                int dim, int accum, uint32_t *albedo, uint32_t *normal, const uint8_t *zero_grids,
                                vct_stats *stats, RasterScratch &rs, hipStream_t s) {
    const size_t nv = (size_t)dim * dim * dim;
    hipError_t e = launch_clear_grids(albedo, normal, dim, zero_grids, s);
    if (e != hipSuccess) return e;
    // the voxelizeInfo buffer is cleared (Application.cpp:581) and this path never counts
    if ((e = launch_fill_u32(&stats->total_voxel_fragments, 0u, 1, s)) != hipSuccess) return e;
    if (sc.ntri == 0) return hipSuccess;
    TessArgs ta;
    ta.dim = dim;
    for (int k = 0; k < 3; ++k) {
        ta.center[k] = center[k];
        ta.vmin[k] = vmin[k];
        ta.vmax[k] = vmax[k];
    }
    hipLaunchKernelGGL(k_tess_setup, dim3((sc.ntri + 255) / 256), dim3(256), 0, s, sc, ta, rs.rec, rs.attr, rs.counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = raster_scan(rs, (size_t)sc.ntri, false, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_tess_points, dim3(kRasterBlocks), dim3(256), 0, s, sc, ta, rs.rec, rs.attr, rs.offs, rs.chunk,
                       rs.chunk_cap, accum, albedo, normal);
    return hipGetLastError();
}

// normalizeVoxels.comp:19-42 for the sum-accumulate mode: per voxel with
// fragments, the averages colour = sum / count (alpha 1, then voxelSetOpacity
// when it is > 0) and normal = sum / count, written to the RGBA8 grids with
// packUnorm4x8, radiance level 0 := (0, 0, 0, alpha), uniqueVoxels and
// maxFragmentsPerVoxel (= the largest count).  Voxels without fragments keep
// the cleared grids and a zero level 0.
__global__ __launch_bounds__(256) void k_normalize_voxels(const uint32_t *__restrict__ sums, uint32_t *albedo,
                                                          uint32_t *normal, int dim, float opacity, LevelInfo L0,
                                                          char *pyr, uint32_t *partials) {
    const size_t nv = (size_t)dim * dim * dim;
    uint32_t uniq = 0, maxc = 0;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < nv; g += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(g % dim), y = (int)((g / dim) % dim), z = (int)(g / ((size_t)dim * dim));
        const uint32_t vi = grid_index(x, y, z, dim >> 2);
        const uint4 c = *(const uint4 *)(sums + (size_t)vi * 8);
        const uint4 n = *(const uint4 *)(sums + (size_t)vi * 8 + 4);
        float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
        if (c.w > 0u) {
            uniq += 1;
            maxc = c.w > maxc ? c.w : maxc;
            const float den = 65535.0f * (float)c.w;
            const float a = opacity > 0.0f ? opacity : 1.0f;
            albedo[vi] = pack_unorm4x8((float)c.x / den, (float)c.y / den, (float)c.z / den, a);
            r.w = a;
        }
        if (n.w > 0u) {
            const float den = 65535.0f * (float)n.w;
            normal[vi] = pack_unorm4x8((float)n.x / den, (float)n.y / den, (float)n.z / den, 1.0f);
        }
        store_stage_texel(pyr, L0, x, y, z, 0, r);
    }
    block_partial(uniq, maxc, 0u, 0u, partials);
}

hipError_t launch_normalize_voxels(const uint32_t *sums, uint32_t *albedo, uint32_t *normal, int dim, float opacity,
                                   const PyrDesc &pyr, vct_stats *stats, uint32_t *partials, hipStream_t s) {
    const int nblk = kStatPartials;
    hipLaunchKernelGGL(k_normalize_voxels, dim3(nblk), dim3(256), 0, s, sums, albedo, normal, dim, opacity, pyr.lv[0],
                       pyr.pyr, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = launch_zero0_reset(pyr, s)) != hipSuccess) return e;  // level 0 written by another pass than the injection
    return launch_reduce_stats(partials, nblk, &stats->unique_voxels, &stats->max_fragments_per_voxel, nullptr, s);
}

}  // namespace vctk
