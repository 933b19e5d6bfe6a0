// radiance.hip -- normalize + inject (transferVoxels.comp, injectRadiance.comp)
// and voxelFillHoles.comp, plus the level <-> linear conversions used by
// download/upload.
//
// The reference injects by scattering one store per shadow-map texel
// (injectRadiance.comp:152-214, radiance[v] := packUnorm4x8(voxelColor[v])).
// Every texel that lands in voxel v stores the same value, so the scatter is
// idempotent and equals a per-voxel "is lit" flag.  Here the scatter only
// stores that flag (one byte per voxel, plain stores: racing writers all write
// 1, so no atomics are needed) and one coalesced per-voxel pass then does
// transferVoxels' work (statistics, opacity rewrite, radiance (0,0,0,a)) and
// the injection together, writing radiance level 0 once in the storage format.
#pragma clang fp contract(off)

#include "vct_device.h"

namespace vctk {

struct InjectLight {
    float ls_inverse[16];
    const float *shadowmap;
    int N, perspective;
};

__global__ __launch_bounds__(256) void k_inject_scatter(InjectLight L, int dim, Extent ext, const float4 *warp,
                                                        uint8_t *litmask) {
    const int i = blockIdx.x * 16 + (threadIdx.x & 15);
    const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= L.N || j >= L.N) return;
    const float tcx = (float)i / (float)L.N, tcy = (float)j / (float)L.N;
    const float d = shadow_bilinear(L.shadowmap, L.N, tcx, tcy, 0, 0);
    float ndc4[4] = {tcx * 2.0f - 1.0f, tcy * 2.0f - 1.0f, d * 2.0f - 1.0f, 1.0f}, w[4];
    mat4_vec4(L.ls_inverse, ndc4, w);
    if (L.perspective) {
        w[0] = w[0] / w[3];
        w[1] = w[1] / w[3];
        w[2] = w[2] / w[3];
    }
    int idx[3];
    float lin[3];
    for (int k = 0; k < 3; ++k) lin[k] = ((w[k] - ext.center[k]) - ext.vmin[k]) / (ext.vmax[k] - ext.vmin[k]);
    if (warp) warp_sample(warp, lin, lin);  // getVoxelPosition with warpTexture (common.glsl:48-50)
    for (int k = 0; k < 3; ++k) idx[k] = (int)((float)dim * lin[k]);
    if (idx[0] < 0 || idx[1] < 0 || idx[2] < 0 || idx[0] >= dim || idx[1] >= dim || idx[2] >= dim) return;
    litmask[((size_t)idx[2] * dim + idx[1]) * dim + idx[0]] = 1;
}

hipError_t launch_inject_scatter(const DevLight *lights, const DevLight *dlights_dev, int nlights,
                                 uint32_t light_mask, int dim, Extent ext, const float4 *warp, uint8_t *litmask,
                                 hipStream_t s) {
    (void)dlights_dev;
    const size_t nv = (size_t)dim * dim * dim;
    hipError_t e = launch_fill_u32((uint32_t *)litmask, 0u, nv / 4, s);
    if (e != hipSuccess) return e;
    for (int li = 0; li < nlights; ++li) {
        const DevLight &L = lights[li];
        if (!((light_mask >> li) & 1u) || !L.shadow_caster || !L.shadowmap) continue;
        InjectLight il;
        for (int k = 0; k < 16; ++k) il.ls_inverse[k] = L.ls_inverse[k];
        il.shadowmap = L.shadowmap;
        il.N = (int)L.shadow_size;
        il.perspective = (int)L.perspective;
        dim3 grid((il.N + 15) / 16, (il.N + 15) / 16);
        hipLaunchKernelGGL(k_inject_scatter, grid, dim3(256), 0, s, il, dim, ext, warp, litmask);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

// One thread per 4-voxel x-row of a brick (16 rows per brick): a wavefront
// covers 4 consecutive bricks, so albedo reads are 64-byte runs and radiance
// writes are whole 512-byte (fp16) bricks.  The grid strides over the rows
// with a fixed number of workgroups, each of which reduces its statistics in
// LDS and writes one partial.
// TEMPORAL (temporalFilterRadiance): the row's previous level-0 values are read
// before they are overwritten, and a voxel that is not lit this frame gets
// mix(previous, (0,0,0,a), 1 - decay) = previous * (1 - w) + (0,0,0,a) * w with
// w = 1 - decay (GLSL mix, transferVoxels.comp:55-58); a lit voxel is
// overwritten with its albedo, (0,0,0,0) if it is empty (injectRadiance.comp:
// 99-101 stores packUnorm4x8(voxelColor) there).
template <int FMT, bool TEMPORAL>
__global__ __launch_bounds__(256) void k_normalize_inject(uint32_t *albedo, const uint8_t *__restrict__ litmask,
                                                          int dim, LevelInfo L0, float opacity, float decay,
                                                          char *pyr, uint32_t *partials) {
    const float mw = 1.0f - decay;  // the mix weight of this frame's value
    const float pw = 1.0f - mw;     // and of the previous one, as GLSL mix computes it
    const int nb = L0.nb;
    const size_t nrows = (size_t)nb * nb * nb * 16;
    uint32_t uniq = 0, maxf = 0, lit_count = 0;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < nrows; g += (size_t)gridDim.x * blockDim.x) {
        const size_t b = g >> 4;
        const int q = (int)(g & 15);
        const int bx = (int)(b % nb), by = (int)((b / nb) % nb), bz = (int)(b / ((size_t)nb * nb));
        const int x0 = bx * 4, y = by * 4 + (q & 3), z = bz * 4 + (q >> 2);
        const size_t li = ((size_t)z * dim + y) * dim + x0;  // the lit mask is linear
        const size_t gi = (b << 6) + (size_t)q * 4;           // the albedo grid is bricked (grid_index)
        const uint4 a = *(const uint4 *)&albedo[gi];
        const uint32_t lw = *(const uint32_t *)&litmask[li];  // 4 flag bytes
        uint32_t w[4] = {a.x, a.y, a.z, a.w};
        float r[4][4];
        bool changed = false;
        char *row = pyr + L0.stage + (size_t)elem_index(L0, x0, y, z, 0) * 16;
        float4 prev[4];
        if (TEMPORAL) {  // the four texels of the row as stored last frame
            if (FMT == 1) {
#pragma unroll
                for (int v = 0; v < 4; ++v) prev[v] = ((const float4 *)row)[v];
            } else {  // fp16 pairs: element x0 holds texels x0, x0+1; element x0+2 texels x0+2, x0+3
                const uint4 e0 = ((const uint4 *)row)[0], e2 = ((const uint4 *)row)[2];
                prev[0] = unpack_h4(e0.x, e0.y);
                prev[1] = unpack_h4(e0.z, e0.w);
                prev[2] = unpack_h4(e2.x, e2.y);
                prev[3] = unpack_h4(e2.z, e2.w);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint32_t a32 = w[v];
            const float ca = unorm8(a32, 3);
            if (ca > 0.0f) {
                uniq += 1;
                const uint32_t mf = (uint32_t)(255.0f * ca);
                maxf = mf > maxf ? mf : maxf;
                const float c0 = unorm8(a32, 0), c1 = unorm8(a32, 1), c2 = unorm8(a32, 2);
                const float na = opacity > 0.0f ? opacity : ca;
                const uint32_t nw = pack_unorm4x8(c0, c1, c2, na);
                if (nw != a32) changed = true;
                w[v] = nw;
                const bool lit = ((lw >> (8 * v)) & 0xffu) != 0u;
                lit_count += lit ? 1u : 0u;
                r[v][0] = lit ? c0 : 0.0f;
                r[v][1] = lit ? c1 : 0.0f;
                r[v][2] = lit ? c2 : 0.0f;
                r[v][3] = na;
                if (TEMPORAL && !lit) {
                    r[v][0] = prev[v].x * pw + 0.0f * mw;
                    r[v][1] = prev[v].y * pw + 0.0f * mw;
                    r[v][2] = prev[v].z * pw + 0.0f * mw;
                    r[v][3] = prev[v].w * pw + na * mw;
                }
            } else {
                r[v][0] = r[v][1] = r[v][2] = r[v][3] = 0.0f;
                if (TEMPORAL && ((lw >> (8 * v)) & 0xffu) == 0u) {
                    r[v][0] = prev[v].x * pw + 0.0f * mw;
                    r[v][1] = prev[v].y * pw + 0.0f * mw;
                    r[v][2] = prev[v].z * pw + 0.0f * mw;
                    r[v][3] = prev[v].w * pw + 0.0f * mw;
                }
            }
        }
        if (changed) *(uint4 *)&albedo[gi] = make_uint4(w[0], w[1], w[2], w[3]);
        // the row's four elements are contiguous (level 0 is isotropic: one face)
        if (FMT == 1) {
            float4 *o = (float4 *)row;
#pragma unroll
            for (int v = 0; v < 4; ++v) o[v] = make_float4(r[v][0], r[v][1], r[v][2], r[v][3]);
        } else {
            // fp16 pairs: elements x0 .. x0+2 complete, the first half of x0+3
            // (its second half is the next row's texel x0+4, written by that
            // row's thread, or the border), and the second half of x0-1
            uint2 h[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) h[v] = pack_h4(make_float4(r[v][0], r[v][1], r[v][2], r[v][3]));
            uint4 *o = (uint4 *)row;
            o[0] = make_uint4(h[0].x, h[0].y, h[1].x, h[1].y);
            o[1] = make_uint4(h[1].x, h[1].y, h[2].x, h[2].y);
            o[2] = make_uint4(h[2].x, h[2].y, h[3].x, h[3].y);
            *(uint2 *)(o + 3) = h[3];
            if (x0 > 0) *(uint2 *)(pyr + L0.stage + (size_t)elem_index(L0, x0 - 1, y, z, 0) * 16 + 8) = h[0];
        }
    }
    block_partial(uniq, maxf, lit_count, 0u, partials);
}

hipError_t launch_normalize_inject(uint32_t *albedo, const uint8_t *litmask, int dim, float opacity, bool temporal,
                                   float decay, const PyrDesc &pyr, vct_stats *stats, uint32_t *partials,
                                   hipStream_t s) {
    const LevelInfo &L0 = pyr.lv[0];
    const size_t nrows = (size_t)L0.nb * L0.nb * L0.nb * 16;
    size_t blocks = (nrows + 255) / 256;
    if (blocks > (size_t)kStatPartials) blocks = kStatPartials;
    dim3 grid((unsigned)blocks);
    if (L0.fmt == 1 && temporal)
        hipLaunchKernelGGL((k_normalize_inject<1, true>), grid, dim3(256), 0, s, albedo, litmask, dim, L0, opacity,
                           decay, pyr.pyr, partials);
    else if (L0.fmt == 1)
        hipLaunchKernelGGL((k_normalize_inject<1, false>), grid, dim3(256), 0, s, albedo, litmask, dim, L0, opacity,
                           decay, pyr.pyr, partials);
    else if (temporal)
        hipLaunchKernelGGL((k_normalize_inject<0, true>), grid, dim3(256), 0, s, albedo, litmask, dim, L0, opacity,
                           decay, pyr.pyr, partials);
    else
        hipLaunchKernelGGL((k_normalize_inject<0, false>), grid, dim3(256), 0, s, albedo, litmask, dim, L0, opacity,
                           decay, pyr.pyr, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // uniqueVoxels, maxFragmentsPerVoxel (transferVoxels.comp:39-45) and the lit count
    return launch_reduce_stats(partials, (int)blocks, &stats->unique_voxels, &stats->max_fragments_per_voxel,
                               &stats->lit_voxels, s);
}

// ---------------------------------------------------------------- fill holes
// voxelFillHoles.comp:8-36 on level 0: src is the level in the pyramid, dst a
// scratch copy with the same layout (copied back afterwards, as the reference
// copies its temporary with glCopyImageSubData).
__global__ __launch_bounds__(256) void k_fill_holes(const char *pyr, char *tmp, LevelInfo L0, int dim) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)dim * dim * dim;
    if (g >= n) return;
    const int x = (int)(g % dim), y = (int)((g / dim) % dim), z = (int)(g / ((size_t)dim * dim));
    float4 cur = load_stage_texel(pyr, L0, x, y, z, 0);
    if (cur.w == 0.0f) {
        float count = 0.0f;
        for (int i = -1; i <= 1; ++i)
            for (int j = -1; j <= 1; ++j)
                for (int k = -1; k <= 1; ++k) {
                    const int xx = x + i, yy = y + j, zz = z + k;
                    if (xx < 0 || yy < 0 || zz < 0 || xx >= dim || yy >= dim || zz >= dim) continue;
                    const float4 v = load_stage_texel(pyr, L0, xx, yy, zz, 0);
                    if (v.w > 0.0f) {
                        cur.x = cur.x + v.x;
                        cur.y = cur.y + v.y;
                        cur.z = cur.z + v.z;
                        cur.w = cur.w + v.w;
                        count += 1.0f;
                    }
                }
        if (count > 0.0f) {
            cur.x = cur.x / count;
            cur.y = cur.y / count;
            cur.z = cur.z / count;
            cur.w = cur.w / count;
        }
    }
    LevelInfo T = L0;
    T.stage = 0;
    store_stage_texel(tmp, T, x, y, z, 0, cur);
}

hipError_t launch_fill_holes(const PyrDesc &pyr, void *tmp, hipStream_t s) {
    const LevelInfo &L0 = pyr.lv[0];
    const size_t n = (size_t)pyr.dim * pyr.dim * pyr.dim;
    hipLaunchKernelGGL(k_fill_holes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pyr.pyr, (char *)tmp, L0,
                       pyr.dim);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t bytes = (size_t)L0.sz * L0.nb * 16;
    return hipMemcpyAsync(pyr.pyr + L0.stage, tmp, bytes, hipMemcpyDeviceToDevice, s);  // glCopyImageSubData
}

// ---------------------------------------------------------------- conversions
__global__ void k_pyr_to_linear(const char *pyr, const float *scales, LevelInfo L, int face, float4 *out) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int size = L.size;
    const size_t n = (size_t)size * size * size;
    if (g >= n) return;
    const int x = (int)(g % size), y = (int)((g / size) % size), z = (int)(g / ((size_t)size * size));
    out[g] = load_texel(pyr, scales, L, x, y, z, face);
}

hipError_t launch_pyr_to_linear(const PyrDesc &pyr, int level, int face, float4 *out, hipStream_t s) {
    const LevelInfo &L = pyr.lv[level];
    const size_t n = (size_t)L.size * L.size * L.size;
    hipLaunchKernelGGL(k_pyr_to_linear, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pyr.pyr, pyr.scales, L,
                       face < L.nfaces ? face : 0, out);
    return hipGetLastError();
}

__global__ void k_linear_to_pyr0(char *pyr, LevelInfo L, const float4 *in) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int size = L.size;
    const size_t n = (size_t)size * size * size;
    if (g >= n) return;
    const int x = (int)(g % size), y = (int)((g / size) % size), z = (int)(g / ((size_t)size * size));
    store_stage_texel(pyr, L, x, y, z, 0, in[g]);
}

hipError_t launch_linear_to_pyr0(const PyrDesc &pyr, const float4 *in, hipStream_t s) {
    const LevelInfo &L = pyr.lv[0];
    const size_t n = (size_t)L.size * L.size * L.size;
    hipLaunchKernelGGL(k_linear_to_pyr0, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pyr.pyr, L, in);
    return hipGetLastError();
}

// bricked voxel grid <-> linear x-fastest (downloads and uploads)
__global__ void k_grid_convert(const uint32_t *src, uint32_t *dst, int dim, int to_linear) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)dim * dim * dim;
    if (g >= n) return;
    const int x = (int)(g % dim), y = (int)((g / dim) % dim), z = (int)(g / ((size_t)dim * dim));
    const uint32_t b = grid_index(x, y, z, dim >> 2);
    if (to_linear) dst[g] = src[b];
    else dst[b] = src[g];
}

hipError_t launch_grid_convert(const uint32_t *src, uint32_t *dst, int dim, bool to_linear, hipStream_t s) {
    const size_t n = (size_t)dim * dim * dim;
    hipLaunchKernelGGL(k_grid_convert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, dim,
                       to_linear ? 1 : 0);
    return hipGetLastError();
}

}  // namespace vctk
