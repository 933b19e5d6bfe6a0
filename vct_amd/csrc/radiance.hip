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

#include <mutex>

#include "vct_device.h"

namespace vctk {

// ---------------------------------------------------------------- no-alloc scratch
// The scratch of vct_normalize_inject_noalloc: the lit mask (one byte per voxel,
// grids up to kNoAllocMaxDim^3) followed by the statistics partials.  Every GPU
// the library runs on holds its own copy until the process ends; every call on
// one GPU shares it, so calls that overlap on one GPU race.
constexpr size_t kNoAllocLitBytes = (size_t)kNoAllocMaxDim * kNoAllocMaxDim * kNoAllocMaxDim;
constexpr size_t kNoAllocScratchBytes = kNoAllocLitBytes + (size_t)kStatPartials * 4 * sizeof(uint32_t);
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kNoAllocMaxDevices];  // device address of `scratch` on each device

// Called by vct_create with `device` current: looks up this device's address of
// the scratch array once (the array's host-side name is not a device address).
hipError_t noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

NoAllocScratch noalloc_scratch(int device) {
    NoAllocScratch s{};
    if (device < 0 || device >= kNoAllocMaxDevices) return s;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    uint8_t *base = (uint8_t *)g_scratch_ptr[device];
    if (!base) return s;
    s.litmask = base;
    s.partials = (uint32_t *)(base + kNoAllocLitBytes);
    return s;
}

struct InjectLight {
    float ls_inverse[16];
    const float *shadowmap;
    int N, perspective;
    double rcp_n;  // RN64(1 / N)
};

// Up to kScatterBatch shadow-casting lights per launch, one per grid z slice.
constexpr int kScatterBatch = 8;
struct InjectBatch {
    InjectLight l[kScatterBatch];
    double rcp_ext[3];  // RN64(1 / (vmax - vmin)) per axis, the difference in float
};

__global__ __launch_bounds__(256) void k_inject_scatter(InjectBatch B, int dim, Extent ext, const float4 *warp,
                                                        uint8_t *litmask) {
    const InjectLight &L = B.l[blockIdx.z];
    const int i = blockIdx.x * 16 + (threadIdx.x & 15);
    const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= L.N || j >= L.N) return;
    // every division is quot_f32 with the divisor's FP64 reciprocal: the IEEE
    // quotient, without a float division sequence per texel
    const float tcx = quot_f32((float)i, L.rcp_n), tcy = quot_f32((float)j, L.rcp_n);
    const float d = shadow_bilinear(L.shadowmap, L.N, tcx, tcy, 0, 0);
    float ndc4[4] = {tcx * 2.0f - 1.0f, tcy * 2.0f - 1.0f, d * 2.0f - 1.0f, 1.0f}, w[4];
    mat4_vec4(L.ls_inverse, ndc4, w);
    if (L.perspective) {
        const double rw = 1.0 / (double)w[3];
        w[0] = quot_f32(w[0], rw);
        w[1] = quot_f32(w[1], rw);
        w[2] = quot_f32(w[2], rw);
    }
    int idx[3];
    float lin[3];
    for (int k = 0; k < 3; ++k) lin[k] = quot_f32((w[k] - ext.center[k]) - ext.vmin[k], B.rcp_ext[k]);
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
    InjectBatch B;
    for (int k = 0; k < 3; ++k) B.rcp_ext[k] = 1.0 / (double)(ext.vmax[k] - ext.vmin[k]);
    int nb = 0, nmax = 0;
    for (int li = 0; li <= nlights; ++li) {
        if (li < nlights) {
            const DevLight &L = lights[li];
            if (!((light_mask >> li) & 1u) || !L.shadow_caster || !L.shadowmap) continue;
            InjectLight &il = B.l[nb++];
            for (int k = 0; k < 16; ++k) il.ls_inverse[k] = L.ls_inverse[k];
            il.shadowmap = L.shadowmap;
            il.N = (int)L.shadow_size;
            il.perspective = (int)L.perspective;
            il.rcp_n = 1.0 / (double)(float)il.N;
            nmax = il.N > nmax ? il.N : nmax;
        }
        if (nb == kScatterBatch || (li == nlights && nb > 0)) {  // each stored texel writes 1: order-free
            dim3 grid((nmax + 15) / 16, (nmax + 15) / 16, nb);
            hipLaunchKernelGGL(k_inject_scatter, grid, dim3(256), 0, s, B, dim, ext, warp, litmask);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            nb = 0;
            nmax = 0;
        }
    }
    return hipSuccess;
}

// Level-0 radiance of one voxel (transferVoxels + injectRadiance): albedo
// word a32 with its opacity rewritten (nw), and the injected value.
// TEMPORAL (temporalFilterRadiance): a voxel that is not lit this frame gets
// mix(previous, (0,0,0,a), 1 - decay) = previous * (1 - w) + (0,0,0,a) * w with
// w = 1 - decay (GLSL mix, transferVoxels.comp:55-58); a lit voxel is
// overwritten with its albedo, (0,0,0,0) if it is empty (injectRadiance.comp:
// 99-101 stores packUnorm4x8(voxelColor) there).
template <bool TEMPORAL>
__device__ __forceinline__ float4 inject_texel(uint32_t a32, bool lit, float opacity, float4 prev, float mw, float pw,
                                               uint32_t *nw) {
    const float ca = unorm8(a32, 3);
    float4 r;
    *nw = a32;
    if (ca > 0.0f) {
        const float c0 = unorm8(a32, 0), c1 = unorm8(a32, 1), c2 = unorm8(a32, 2);
        const float na = opacity > 0.0f ? opacity : ca;
        *nw = pack_unorm4x8(c0, c1, c2, na);
        r = make_float4(lit ? c0 : 0.0f, lit ? c1 : 0.0f, lit ? c2 : 0.0f, na);
        if (TEMPORAL && !lit)
            r = make_float4(prev.x * pw + 0.0f * mw, prev.y * pw + 0.0f * mw, prev.z * pw + 0.0f * mw,
                            prev.w * pw + na * mw);
    } else {
        r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (TEMPORAL && !lit)
            r = make_float4(prev.x * pw + 0.0f * mw, prev.y * pw + 0.0f * mw, prev.z * pw + 0.0f * mw,
                            prev.w * pw + 0.0f * mw);
    }
    return r;
}

// The same texel in the fp16 storage format without the temporal filter,
// from tables instead of arithmetic: h16[k] holds the fp16 bits of k/255 (the
// unorm8 value, correctly rounded to float, then to fp16, as pack_h4 of
// inject_texel gives them).  The rewritten albedo word is the old word with
// its alpha byte replaced by the opacity's: packUnorm4x8 of k/255 gives back
// k for every byte (checked exhaustively), so the colour bytes are unchanged.
// op_b / op_h: the opacity as an unorm8 byte and as fp16 bits (opacity > 0).
__device__ __forceinline__ uint2 inject_texel_h(uint32_t a32, bool lit, float opacity, uint32_t op_b, uint32_t op_h,
                                                const uint16_t *h16, uint32_t *nw) {
    const uint32_t ab = a32 >> 24;
    *nw = a32;
    if (ab == 0u) return make_uint2(0u, 0u);
    uint32_t na_h = op_h;
    if (opacity > 0.0f) *nw = (a32 & 0x00ffffffu) | (op_b << 24);
    else na_h = h16[ab];
    if (!lit) return make_uint2(0u, na_h << 16);
    return make_uint2((uint32_t)h16[a32 & 0xffu] | ((uint32_t)h16[(a32 >> 8) & 0xffu] << 16),
                      (uint32_t)h16[(a32 >> 16) & 0xffu] | (na_h << 16));
}

// One wavefront per 4^3 brick, one voxel per lane in the brick's element
// order (lane = (x&3) | (y&3) << 2 | (z&3) << 4, the same for the bricked
// albedo grid and for level 0), so albedo reads are one 256-byte run and each
// level-0 store is one contiguous run of the brick's elements: 1 KB of fp16
// pairs or fp32 texels.  An fp16 pair element x holds texels x and x+1: the
// second comes from the next lane, or for the last voxel of a brick row from
// that voxel's own inputs in the next brick (texel dim is the zero border).
// The grid strides over the bricks with a fixed number of workgroups, each of
// which reduces its statistics in LDS and writes one partial.
// The opacity rewrite of the albedo word (rewrite = 1) races with the read of
// that word by the neighbouring brick's wavefront; either value gives the same
// texel, except when the opacity rounds to an alpha byte of 0 -- then the
// launcher passes rewrite = 0 and rewrites in a pass of its own afterwards.
template <int FMT, bool TEMPORAL>
__global__ __launch_bounds__(256) void k_normalize_inject(uint32_t *albedo, const uint8_t *__restrict__ litmask,
                                                          int dim, LevelInfo L0, float opacity, float decay,
                                                          int rewrite, char *pyr, uint8_t *zero0,
                                                          uint32_t *partials) {
    const float mw = 1.0f - decay;  // the mix weight of this frame's value
    const float pw = 1.0f - mw;     // and of the previous one, as GLSL mix computes it
    // per byte k: the fp16 bits of k/255, and maxFragmentsPerVoxel's
    // uint(255 * (k/255)) (transferVoxels.comp:41) as the float product gives it
    __shared__ uint16_t h16[256];
    __shared__ uint8_t mf8[256];
    {
        const float v = unorm8((uint32_t)threadIdx.x, 0);  // 256 threads, one byte each
        h16[threadIdx.x] = f2h(v);
        mf8[threadIdx.x] = (uint8_t)(uint32_t)(255.0f * v);
    }
    __syncthreads();
    const uint32_t op_b = pack_unorm4x8(0.0f, 0.0f, 0.0f, opacity) >> 24;
    const uint32_t op_h = (uint32_t)f2h(opacity);
    const int nb = L0.nb;
    const uint32_t nbricks = (uint32_t)nb * nb * nb;
    const int lgnb = __builtin_ctz((unsigned)nb);
    const int lane = threadIdx.x & 63;
    const int vx = lane & 3, vy = (lane >> 2) & 3, vz = lane >> 4;
    uint32_t uniq = 0, maxf = 0, lit_count = 0;
    for (uint32_t b = blockIdx.x * 4u + (threadIdx.x >> 6); b < nbricks; b += gridDim.x * 4u) {
        // nb = dim / 4 is a power of two (dim is one): shifts and masks
        const int bx = (int)(b & (uint32_t)(nb - 1)), by = (int)((b >> lgnb) & (uint32_t)(nb - 1)),
                  bz = (int)(b >> (2 * lgnb));
        const int x = bx * 4 + vx, y = by * 4 + vy, z = bz * 4 + vz;
        const size_t gi = ((size_t)b << 6) + (size_t)lane;  // grid_index(x, y, z)
        const size_t li = ((size_t)z * dim + y) * dim + x;   // the lit mask is linear
        const uint32_t a32 = albedo[gi];
        const bool lit = litmask[li] != 0;
        char *elem = pyr + L0.stage + (size_t)elem_index(L0, x, y, z, 0) * 16;
        // the brick's elements already hold +0 (wave-uniform)
        const uint32_t z0 = zero0 != nullptr ? (uint32_t)zero0[b] : 0u;  // PyrDesc::zero0 bits
        const bool was0 = (z0 & 1u) != 0u;
        float4 prev = make_float4(0.f, 0.f, 0.f, 0.f), prev1 = prev;
        if (TEMPORAL && !was0) {  // this element as stored last frame (texels x, x+1 for fp16 pairs)
            const uint4 e = *(const uint4 *)elem;
            if (FMT == 1) {
                prev = make_float4(__uint_as_float(e.x), __uint_as_float(e.y), __uint_as_float(e.z),
                                   __uint_as_float(e.w));
            } else {
                prev = unpack_h4(e.x, e.y);
                prev1 = unpack_h4(e.z, e.w);
            }
        }
        uint32_t nw;
        constexpr bool LUT = FMT == 0 && !TEMPORAL;
        float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
        uint2 h = make_uint2(0u, 0u);
        if (LUT) h = inject_texel_h(a32, lit, opacity, op_b, op_h, h16, &nw);
        else r = inject_texel<TEMPORAL>(a32, lit, opacity, prev, mw, pw, &nw);
        if ((a32 >> 24) != 0u) {
            uniq += 1;
            const uint32_t mf = mf8[a32 >> 24];
            maxf = mf > maxf ? mf : maxf;
            lit_count += lit ? 1u : 0u;
        }
        if (rewrite && nw != a32) albedo[gi] = nw;
        // every albedo word of the brick is 0 (a word that is 0 stays 0 under
        // any later opacity rewrite); and was 0 as the voxelizer left it
        const uint32_t za = (__ballot(nw != 0u) == 0 ? 2u : 0u) | (__ballot(a32 != 0u) == 0 ? 4u : 0u);
        if (FMT == 1) {
            const bool nz = (__float_as_uint(r.x) | __float_as_uint(r.y) | __float_as_uint(r.z) |
                             __float_as_uint(r.w)) != 0u;
            const bool now0 = __ballot(nz) == 0;
            if (!(now0 && was0)) *(float4 *)elem = r;  // an all-zero brick over zeros is not rewritten
            const uint32_t z1 = (now0 ? 1u : 0u) | za;
            if (zero0 != nullptr && lane == 0 && z1 != z0) zero0[b] = (uint8_t)z1;
            continue;
        }
        if (!LUT) h = pack_h4(r);
        uint2 h1 = make_uint2(__shfl_down(h.x, 1, 64), __shfl_down(h.y, 1, 64));  // texel x+1 in lane + 1
        if (vx == 3) {
            h1 = make_uint2(0u, 0u);  // texel dim: the border
            if (x + 1 < dim) {        // voxel x+1 opens brick b+1 along x
                const uint32_t a1 = albedo[gi + 64 - 3];
                const bool lit1 = litmask[li + 1] != 0;
                uint32_t nw1;
                if (LUT) h1 = inject_texel_h(a1, lit1, opacity, op_b, op_h, h16, &nw1);
                else h1 = pack_h4(inject_texel<TEMPORAL>(a1, lit1, opacity, prev1, mw, pw, &nw1));
            }
        }
        const bool now0 = __ballot((h.x | h.y | h1.x | h1.y) != 0u) == 0;
        if (!(now0 && was0)) {  // an all-zero brick over zeros is not rewritten
            *(uint4 *)elem = make_uint4(h.x, h.y, h1.x, h1.y);
            if (x == 0)  // texel 0 is also the second half of element -1, in the border brick
                *(uint2 *)(pyr + (uint32_t)(L0.stage + elem_index(L0, -1, y, z, 0) * 16u) + 8) = h;
        }
        const uint32_t z1 = (now0 ? 1u : 0u) | za;
        if (zero0 != nullptr && lane == 0 && z1 != z0) zero0[b] = (uint8_t)z1;
    }
    block_partial(uniq, maxf, lit_count, 0u, partials);
}

// The opacity rewrite on its own (transferVoxels.comp:47-50), after the
// injection, for opacities whose alpha byte is 0.
__global__ __launch_bounds__(256) void k_rewrite_opacity(uint32_t *albedo, size_t n, float opacity) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t a32 = albedo[i];
        uint32_t nw;
        inject_texel<false>(a32, false, opacity, make_float4(0.f, 0.f, 0.f, 0.f), 0.f, 0.f, &nw);
        if (nw != a32) albedo[i] = nw;
    }
}

hipError_t launch_normalize_inject(uint32_t *albedo, const uint8_t *litmask, int dim, float opacity, bool temporal,
                                   float decay, const PyrDesc &pyr, vct_stats *stats, uint32_t *partials,
                                   hipStream_t s) {
    const LevelInfo &L0 = pyr.lv[0];
    const size_t nbricks = (size_t)L0.nb * L0.nb * L0.nb;
    size_t blocks = (nbricks + 3) / 4;
    if (blocks > (size_t)kStatPartials) blocks = kStatPartials;
    dim3 grid((unsigned)blocks);
    // pack_unorm4x8 rounds opacity * 255: below 0.5 the rewritten alpha byte is 0
    const int rewrite = (opacity > 0.0f && opacity * 255.0f < 0.5f) ? 0 : 1;
#define VCT_INJECT_LAUNCH(F, T)                                                                                     \
    hipLaunchKernelGGL((k_normalize_inject<F, T>), grid, dim3(256), 0, s, albedo, litmask, dim, L0, opacity, decay, \
                       rewrite, pyr.pyr, pyr.zero0, partials)
    if (L0.fmt == 1 && temporal) VCT_INJECT_LAUNCH(1, true);
    else if (L0.fmt == 1) VCT_INJECT_LAUNCH(1, false);
    else if (temporal) VCT_INJECT_LAUNCH(0, true);
    else VCT_INJECT_LAUNCH(0, false);
#undef VCT_INJECT_LAUNCH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (!rewrite) {
        const size_t nv = (size_t)dim * dim * dim;
        hipLaunchKernelGGL(k_rewrite_opacity, dim3(2048), dim3(256), 0, s, albedo, nv, opacity);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
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
    T.stage = L0.pad * 16u;  // tmp holds the stored level, border bricks included
    store_stage_texel(tmp, T, x, y, z, 0, cur);
}

hipError_t launch_fill_holes(const PyrDesc &pyr, void *tmp, hipStream_t s) {
    const LevelInfo &L0 = pyr.lv[0];
    const size_t n = (size_t)pyr.dim * pyr.dim * pyr.dim;
    hipLaunchKernelGGL(k_fill_holes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pyr.pyr, (char *)tmp, L0,
                       pyr.dim);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t bytes = (size_t)L0.nelem * 16;
    e = hipMemcpyAsync(pyr.pyr + L0.stage - (size_t)L0.pad * 16, tmp, bytes, hipMemcpyDeviceToDevice,
                       s);  // glCopyImageSubData
    if (e != hipSuccess) return e;
    return launch_zero0_reset(pyr, s);
}

hipError_t launch_zero0_reset(const PyrDesc &pyr, hipStream_t s) {
    if (!pyr.zero0) return hipSuccess;
    const size_t nb = (size_t)pyr.lv[0].nb;
    return launch_fill_u32((uint32_t *)pyr.zero0, 0u, (nb * nb * nb + 3) / 4, s);
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
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_zero0_reset(pyr, s);
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
