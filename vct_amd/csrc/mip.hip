// mip.hip -- the radiance mip chain (filterRadiance.comp KERNEL_BOX2,
// src/Application.cpp:877-921) fused over three levels per launch.
//
// The reference dispatches one compute pass per level with a full barrier in
// between, re-reading each level from memory.  Here one workgroup loads an
// 8^3 region of the source level (two 4^3 bricks per axis, contiguous 512-byte
// brick runs) into LDS and produces the 4^3, 2^3 and 1^3 blocks of the next
// three levels from LDS, so level 0 -- the only large one -- is read exactly
// once.  Each level is rounded to its storage format before the next one is
// filtered from it, exactly as a per-level pass over stored texels would be,
// and the eight taps are summed in the reference's offset order, so the chain
// is bit-identical to filtering level by level.
//
// Anisotropic chains keep six directional volumes per level >= 1 (+x,-x,+y,
// -y,+z,-z): each is the mean of the four front-to-back composites
// near + (1 - near.a) * far of the child pairs along that axis.
#pragma clang fp contract(off)

#include "vct_device.h"

namespace vctk {

struct MipArgs {
    char *pyr;
    LevelInfo src;
    LevelInfo out[3];
    int nout;
    int fp32;  // stage format: 1 = fp32x4, 0 = fp16 pairs
    // source level 0 only: its bricks known to hold +0 (PyrDesc::zero0), or null
    const uint8_t *zero0;
};

// Whether brick (x, y, z) >> 2 of the source is known to hold +0: its texels
// are then not read.
__device__ __forceinline__ bool src_zero(const MipArgs &a, int x, int y, int z) {
    if (a.zero0 == nullptr) return false;
    const uint32_t nb = (uint32_t)a.src.nb;
    return (a.zero0[((uint32_t)(z >> 2) * nb + (uint32_t)(y >> 2)) * nb + (uint32_t)(x >> 2)] & 1u) != 0;
}

__device__ __forceinline__ float4 round_stage(float4 v, int fp32) {
    if (fp32) return v;
    return make_float4(q16(v.x), q16(v.y), q16(v.z), q16(v.w));
}

// One filtering step of the chain: the (cur/2)^3 texels per face of output
// level k from the cur^3 texels per face in LDS (faces: 1 or 6 source faces).
// CUR: cur as a compile-time constant (0: use the argument), which turns the
// index arithmetic into shifts.
template <bool ANISO, int CUR>
__device__ __forceinline__ void mip_step(const MipArgs &a, int cur_rt, int faces, const float4 *in, int in_stride,
                                         float4 *out, int out_stride, int k, int ox, int oy, int oz) {
    constexpr int NF = ANISO ? 6 : 1;
    const int cur = CUR ? CUR : cur_rt;
    const int d = cur >> 1;
    const LevelInfo &ol = a.out[k];
    const int sh = k + 1;
    const int gx = ox >> sh, gy = oy >> sh, gz = oz >> sh;
    const int d3 = d * d * d;
    for (int i = threadIdx.x; i < d3 * NF; i += 256) {
        const int f = i / d3;
        const int vi = i - f * d3;
        const int x = vi % d, y = (vi / d) % d, z = vi / (d * d);
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!ANISO) {
            // BOX2 offsets in filterRadiance.comp order: (0,0,0),(0,0,1),(0,1,0),(0,1,1),(1,0,0),...
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                const int dx = (o >> 2) & 1, dy = (o >> 1) & 1, dz = o & 1;
                const float4 t = in[((2 * z + dz) * cur + (2 * y + dy)) * cur + (2 * x + dx)];
                v.x = v.x + t.x;
                v.y = v.y + t.y;
                v.z = v.z + t.z;
                v.w = v.w + t.w;
            }
            v.x = v.x * 0.125f;
            v.y = v.y * 0.125f;
            v.z = v.z * 0.125f;
            v.w = v.w * 0.125f;
        } else {
            const int ax = f >> 1, neg = f & 1;
            const int sf = faces == 1 ? 0 : f;
            const float4 *src = in + (size_t)sf * in_stride;
            const int oa = ax == 0 ? 1 : 0, ob = ax == 2 ? 1 : 2;
#pragma unroll
            for (int pa = 0; pa < 2; ++pa)
#pragma unroll
                for (int pb = 0; pb < 2; ++pb) {
                    int cn[3] = {2 * x, 2 * y, 2 * z}, cf[3] = {2 * x, 2 * y, 2 * z};
                    cn[oa] += pa;
                    cf[oa] += pa;
                    cn[ob] += pb;
                    cf[ob] += pb;
                    cn[ax] += neg ? 1 : 0;
                    cf[ax] += neg ? 0 : 1;
                    const float4 pn = src[(cn[2] * cur + cn[1]) * cur + cn[0]];
                    const float4 pf = src[(cf[2] * cur + cf[1]) * cur + cf[0]];
                    const float t = 1.0f - pn.w;
                    v.x = v.x + (pn.x + t * pf.x);
                    v.y = v.y + (pn.y + t * pf.y);
                    v.z = v.z + (pn.z + t * pf.z);
                    v.w = v.w + (pn.w + t * pf.w);
                }
            v.x = v.x * 0.25f;
            v.y = v.y * 0.25f;
            v.z = v.z * 0.25f;
            v.w = v.w * 0.25f;
        }
        v = round_stage(v, a.fp32);
        out[(size_t)f * out_stride + vi] = v;
        store_stage_texel(a.pyr, ol, gx + x, gy + y, gz + z, NF == 6 ? f : 0, v);
    }
}

// The first step of an 8^3 region (4^3 texels per face): wavefront w filters
// faces w and w + 4, one texel per lane.  The face is wave-uniform, so the
// child offsets of its near/far pairs are scalar values and only the texel's
// base address is per lane.  Same taps, in the same order, as mip_step.
template <bool ANISO, bool SRC6>
__device__ __forceinline__ void mip_step8(const MipArgs &a, const float4 *in, float4 *out, int ox, int oy, int oz) {
    constexpr int NF = ANISO ? 6 : 1;
    constexpr int cur = 8;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int x = lane & 3, y = (lane >> 2) & 3, z = lane >> 4;
    const int base = ((2 * z) * cur + 2 * y) * cur + 2 * x;
    const LevelInfo &ol = a.out[0];
    for (int f = w; f < NF; f += 4) {
        const float4 *c = in + (SRC6 ? f * 512 : 0) + base;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!ANISO) {
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                const int dx = (o >> 2) & 1, dy = (o >> 1) & 1, dz = o & 1;
                const float4 t = c[(dz * cur + dy) * cur + dx];
                v.x = v.x + t.x;
                v.y = v.y + t.y;
                v.z = v.z + t.z;
                v.w = v.w + t.w;
            }
            v.x = v.x * 0.125f;
            v.y = v.y * 0.125f;
            v.z = v.z * 0.125f;
            v.w = v.w * 0.125f;
        } else {
            const int ax = f >> 1, neg = f & 1;
            // strides of the face's axis and of the two axes across it (oa, ob
            // as in mip_step: y, z for x; x, z for y; x, y for z)
            const int s_ax = ax == 0 ? 1 : (ax == 1 ? cur : cur * cur);
            const int s_oa = ax == 0 ? cur : 1;
            const int s_ob = ax == 2 ? cur : cur * cur;
            const int near = neg ? s_ax : 0, far = neg ? 0 : s_ax;
#pragma unroll
            for (int pa = 0; pa < 2; ++pa)
#pragma unroll
                for (int pb = 0; pb < 2; ++pb) {
                    const int o = pa * s_oa + pb * s_ob;
                    const float4 pn = c[o + near];
                    const float4 pf = c[o + far];
                    const float t = 1.0f - pn.w;
                    v.x = v.x + (pn.x + t * pf.x);
                    v.y = v.y + (pn.y + t * pf.y);
                    v.z = v.z + (pn.z + t * pf.z);
                    v.w = v.w + (pn.w + t * pf.w);
                }
            v.x = v.x * 0.25f;
            v.y = v.y * 0.25f;
            v.z = v.z * 0.25f;
            v.w = v.w * 0.25f;
        }
        v = round_stage(v, a.fp32);
        out[f * 64 + lane] = v;
        store_stage_texel(a.pyr, ol, (ox >> 1) + x, (oy >> 1) + y, (oz >> 1) + z, NF == 6 ? f : 0, v);
    }
}

// ANISO: outputs have six faces; SRC6: the source level has six faces (any
// anisotropic level >= 1).  LDS: the 8^3 source region (per face), then the
// ping-pong buffers of the three output steps (4^3, 2^3, 1^3 per face).
template <bool ANISO, bool SRC6>
__global__ __launch_bounds__(256) void k_mip_chain(MipArgs a) {
    constexpr int NF = ANISO ? 6 : 1;
    constexpr int SF = SRC6 ? 6 : 1;
    __shared__ float4 bufA[SF][512];
    __shared__ float4 bufB[NF][64];
    const int tid = threadIdx.x;
    const int s = a.src.size;
    const int R = s < 8 ? s : 8;
    const int nreg = s / R;
    // XCD-aware order: workgroup b runs on XCD b % 8, so give each XCD a
    // contiguous run of regions; x-neighbours then share an L2 and run at
    // about the same time, which merges the two halves of the pair elements
    // on a region's x boundary in one L2 line before it is written back.
    const int nblk = nreg * nreg * nreg;
    const int b = (nblk & 7) == 0 ? (int)(blockIdx.x & 7u) * (nblk >> 3) + (int)(blockIdx.x >> 3) : (int)blockIdx.x;
    const int rx = b % nreg, ry = (b / nreg) % nreg, rz = b / (nreg * nreg);
    const int ox = rx * R, oy = ry * R, oz = rz * R;
    const int R3 = R * R * R;
    for (int f = 0; f < SF; ++f) {  // the launcher picks SRC6 from the source level's face count
        if (R == 8 && !a.fp32) {
            // fp16 pairs: the elements of even x hold texels x and x+1; walk
            // brick by brick (32 such elements per brick)
            const int i = tid;  // 256 threads, 256 pair elements
            const int j = i >> 5, v = i & 31;
            const int x = ox + (j & 1) * 4 + (v & 1) * 2;
            const int y = oy + ((j >> 1) & 1) * 4 + ((v >> 1) & 3);
            const int z = oz + (j >> 2) * 4 + (v >> 3);
            uint4 e = make_uint4(0u, 0u, 0u, 0u);
            if (!src_zero(a, x, y, z)) e = *(const uint4 *)(a.pyr + a.src.stage + (size_t)elem_index(a.src, x, y, z, f) * 16);
            float4 *d = &bufA[f][((z - oz) * R + (y - oy)) * R + (x - ox)];
            d[0] = unpack_h4(e.x, e.y);
            d[1] = unpack_h4(e.z, e.w);
        } else {
            for (int i = tid; i < R3; i += 256) {
                int x, y, z;
                if (R == 8) {  // brick-major walk: 8 bricks of 64 contiguous voxels
                    const int j = i >> 6, v = i & 63;
                    x = ox + (j & 1) * 4 + (v & 3);
                    y = oy + ((j >> 1) & 1) * 4 + ((v >> 2) & 3);
                    z = oz + (j >> 2) * 4 + (v >> 4);
                } else {
                    x = i % R;
                    y = (i / R) % R;
                    z = i / (R * R);
                }
                bufA[f][((z - oz) * R + (y - oy)) * R + (x - ox)] =
                    src_zero(a, x, y, z) ? make_float4(0.f, 0.f, 0.f, 0.f) : load_stage_texel(a.pyr, a.src, x, y, z, f);
            }
        }
    }
    __syncthreads();
    if (R == 8) {  // the three steps with their sizes known at compile time
        mip_step8<ANISO, SRC6>(a, &bufA[0][0], &bufB[0][0], ox, oy, oz);
        if (a.nout > 1) {
            __syncthreads();
            mip_step<ANISO, 4>(a, 4, NF, &bufB[0][0], 64, &bufA[0][0], 8, 1, ox, oy, oz);
        }
        if (a.nout > 2) {
            __syncthreads();
            mip_step<ANISO, 2>(a, 2, NF, &bufA[0][0], 8, &bufB[0][0], 64, 2, ox, oy, oz);
        }
        return;
    }
    int cur = R, faces = SF;
    float4 *in = &bufA[0][0];
    int in_stride = 512;
    for (int k = 0; k < a.nout; ++k) {
        const int d = cur >> 1;
        if (d < 1) break;
        // ping-pong: outputs of step k go to bufB (k even, <= 4^3 per face) or
        // bufA (k odd, <= 2^3 per face)
        float4 *out = (k & 1) ? &bufA[0][0] : &bufB[0][0];
        const int out_stride = (k & 1) ? 8 : 64;
        mip_step<ANISO, 0>(a, cur, faces, in, in_stride, out, out_stride, k, ox, oy, oz);
        __syncthreads();
        in = out;
        in_stride = out_stride;
        cur = d;
        faces = NF;
    }
}

// FP8 E4M3 encode of one level: one 64-thread wavefront per (brick, face);
// the scale is the largest magnitude in the brick over 448 (1 if all zero).
__global__ __launch_bounds__(64) void k_fp8_encode(char *pyr, float *scales, LevelInfo L) {
    const int bf = blockIdx.x;  // brick * nfaces + face
    const int b = bf / L.nfaces, f = bf % L.nfaces;
    const int v = threadIdx.x;
    const int nb = L.nb;
    const int bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
    const int x = bx * 4 + (v & 3), y = by * 4 + ((v >> 2) & 3), z = bz * 4 + (v >> 4);
    const bool valid = x < L.size && y < L.size && z < L.size;
    float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid) c = load_stage_texel(pyr, L, x, y, z, f);
    float m = fmaxf(fmaxf(__builtin_fabsf(c.x), __builtin_fabsf(c.y)), fmaxf(__builtin_fabsf(c.z), __builtin_fabsf(c.w)));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    const float sc = m > 0.0f ? m / 448.0f : 1.0f;
    const uint32_t e = elem_index(L, x, y, z, f);  // brick and face: e / 64 (every stride is a multiple of 64)
    if (v == 0) scales[L.scale + (e >> 6)] = sc;
    const uint32_t code = f2fp8(c.x / sc) | (f2fp8(c.y / sc) << 8) | (f2fp8(c.z / sc) << 16) | (f2fp8(c.w / sc) << 24);
    *(uint32_t *)(pyr + L.base + (size_t)e * 4) = code;
}

// ---------------------------------------------------------------- albedo chain
// The voxelColor chain of the same mipmap pass (Application.cpp:904-917): BOX2
// on an RGBA8 image, so every level is stored as packUnorm4x8 of the mean of
// its eight children read back as k/255 (filterRadiance.comp:14-65 with
// rgba8 image bindings).  Levels >= 1 are linear, x fastest.  Levels 1 and 2
// come from k_albedo_bricks (below); the coarser ones from k_albedo_chain,
// the region scheme of k_mip_chain: a workgroup loads an 8^3 region of a
// linear source level and writes up to three levels from LDS.
struct AlbedoMipArgs {
    const uint32_t *src;
    int src_size;
    uint32_t *out[3];
    int nout;
};

__global__ __launch_bounds__(256) void k_albedo_chain(AlbedoMipArgs a) {
    __shared__ float4 bufA[512];
    __shared__ float4 bufB[64];
    const int tid = threadIdx.x;
    const int s = a.src_size;
    const int R = s < 8 ? s : 8;
    const int nreg = s / R;
    const int b = (int)blockIdx.x;
    const int ox = (b % nreg) * R, oy = ((b / nreg) % nreg) * R, oz = (b / (nreg * nreg)) * R;
    const int R3 = R * R * R;
    for (int i = tid; i < R3; i += 256) {
        const int x = ox + i % R, y = oy + (i / R) % R, z = oz + i / (R * R);
        const uint32_t w = a.src[((size_t)z * s + y) * s + x];
        bufA[((z - oz) * R + (y - oy)) * R + (x - ox)] = make_float4(unorm8(w, 0), unorm8(w, 1), unorm8(w, 2),
                                                                     unorm8(w, 3));
    }
    __syncthreads();
    int cur = R;
    float4 *in = bufA;
    for (int k = 0; k < a.nout; ++k) {
        const int d = cur >> 1;
        if (d < 1) break;
        float4 *out = (k & 1) ? bufA : bufB;
        const int os = s >> (k + 1);  // output level size
        const int gx = ox >> (k + 1), gy = oy >> (k + 1), gz = oz >> (k + 1);
        for (int i = tid; i < d * d * d; i += 256) {
            const int x = i % d, y = (i / d) % d, z = i / (d * d);
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            // BOX2 offsets in filterRadiance.comp order: (0,0,0),(0,0,1),(0,1,0),(0,1,1),(1,0,0),...
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                const int dx = (o >> 2) & 1, dy = (o >> 1) & 1, dz = o & 1;
                const float4 t = in[((2 * z + dz) * cur + (2 * y + dy)) * cur + (2 * x + dx)];
                v.x = v.x + t.x;
                v.y = v.y + t.y;
                v.z = v.z + t.z;
                v.w = v.w + t.w;
            }
            const uint32_t pw = pack_unorm4x8(v.x * 0.125f, v.y * 0.125f, v.z * 0.125f, v.w * 0.125f);
            a.out[k][((size_t)(gz + z) * os + (gy + y)) * os + (gx + x)] = pw;
            out[i] = make_float4(unorm8(pw, 0), unorm8(pw, 1), unorm8(pw, 2), unorm8(pw, 3));
        }
        __syncthreads();
        in = out;
        cur = d;
    }
}

// Levels 1 and 2 of the albedo chain straight from the bricked grid: one
// thread per 4^3 brick, which holds exactly the 2^3 level-1 texels and the one
// level-2 texel it filters to.  The brick is read a z-slab pair at a time
// (eight 16-byte loads of 32 words), every level-1 texel is rounded to RGBA8
// before level 2 is filtered from it, and the taps are summed in the BOX2
// order of filterRadiance.comp -- the same values as filtering level by level.
__global__ __launch_bounds__(256) void k_albedo_bricks(const uint32_t *__restrict__ grid,
                                                       const uint8_t *__restrict__ zero0, int dim, uint32_t *out1,
                                                       uint32_t *out2) {
    const int nb = dim >> 2;
    const uint32_t b = blockIdx.x * 256u + threadIdx.x;
    if (b >= (uint32_t)nb * nb * nb) return;
    const int lg = __builtin_ctz((unsigned)nb);
    const int bx = (int)(b & (uint32_t)(nb - 1)), by = (int)((b >> lg) & (uint32_t)(nb - 1)), bz = (int)(b >> (2 * lg));
    if (zero0 != nullptr && (zero0[b] & 2u) != 0) {  // bit 1: every word of the brick is 0
        // every word of the brick is 0: the texels it filters to are
        // packUnorm4x8 of sums of zeros, 0, and the brick is not read
        const int s1 = dim >> 1;
#pragma unroll
        for (int cz = 0; cz < 2; ++cz)
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
                *(uint2 *)&out1[((size_t)(bz * 2 + cz) * s1 + (by * 2 + cy)) * s1 + bx * 2] = make_uint2(0u, 0u);
        if (out2 != nullptr) {
            const int s2 = dim >> 2;
            out2[((size_t)bz * s2 + by) * s2 + bx] = 0u;
        }
        return;
    }
    const uint4 *src = (const uint4 *)(grid + ((size_t)b << 6));
    const int s1 = dim >> 1;
    float4 l1[8];  // the rounded level-1 texels, (cx, cy, cz) = bits (2, 1, 0) in BOX2 order
#pragma unroll
    for (int cz = 0; cz < 2; ++cz) {
        uint32_t w[32];  // z rows 2cz, 2cz+1 of the brick: w[(z & 1) * 16 + y * 4 + x]
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint4 v = src[cz * 8 + k];
            w[k * 4 + 0] = v.x;
            w[k * 4 + 1] = v.y;
            w[k * 4 + 2] = v.z;
            w[k * 4 + 3] = v.w;
        }
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cx = 0; cx < 2; ++cx) {
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int o = 0; o < 8; ++o) {  // (0,0,0),(0,0,1),(0,1,0),(0,1,1),(1,0,0),...
                    const int dx = (o >> 2) & 1, dy = (o >> 1) & 1, dz = o & 1;
                    const uint32_t t = w[dz * 16 + (2 * cy + dy) * 4 + (2 * cx + dx)];
                    v.x = v.x + unorm8(t, 0);
                    v.y = v.y + unorm8(t, 1);
                    v.z = v.z + unorm8(t, 2);
                    v.w = v.w + unorm8(t, 3);
                }
                const uint32_t pw = pack_unorm4x8(v.x * 0.125f, v.y * 0.125f, v.z * 0.125f, v.w * 0.125f);
                const int x = bx * 2 + cx, y = by * 2 + cy, z = bz * 2 + cz;
                out1[((size_t)z * s1 + y) * s1 + x] = pw;
                l1[(cx << 2) | (cy << 1) | cz] = make_float4(unorm8(pw, 0), unorm8(pw, 1), unorm8(pw, 2), unorm8(pw, 3));
            }
    }
    if (out2 == nullptr) return;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        v.x = v.x + l1[o].x;
        v.y = v.y + l1[o].y;
        v.z = v.z + l1[o].z;
        v.w = v.w + l1[o].w;
    }
    const int s2 = dim >> 2;
    out2[((size_t)bz * s2 + by) * s2 + bx] = pack_unorm4x8(v.x * 0.125f, v.y * 0.125f, v.z * 0.125f, v.w * 0.125f);
}

hipError_t launch_albedo_chain(const uint32_t *grid, const uint8_t *zero0, uint32_t *const *levels, int dim, int nlevels,
                               hipStream_t s) {
    if (nlevels < 2) return hipSuccess;
    {
        const unsigned nbricks = (unsigned)((dim >> 2) * (dim >> 2) * (dim >> 2));
        hipLaunchKernelGGL(k_albedo_bricks, dim3((nbricks + 255) / 256), dim3(256), 0, s, grid, zero0, dim, levels[1],
                           nlevels > 2 ? levels[2] : nullptr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    int src = nlevels > 2 ? 2 : 1;
    while (src + 1 < nlevels) {
        const int size = dim >> src;
        const int R = size < 8 ? size : 8;
        if (R < 2) break;
        const int lg = R == 8 ? 3 : (R == 4 ? 2 : 1);
        int nout = nlevels - 1 - src;
        if (nout > lg) nout = lg;
        AlbedoMipArgs a;
        a.src = levels[src];
        a.src_size = size;
        for (int k = 0; k < 3; ++k) a.out[k] = levels[k < nout ? src + 1 + k : src + 1];
        a.nout = nout;
        const int nreg = size / R;
        hipLaunchKernelGGL(k_albedo_chain, dim3((unsigned)(nreg * nreg * nreg)), dim3(256), 0, s, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        src += nout;
    }
    return hipSuccess;
}

hipError_t launch_mip(const PyrDesc &pyr, hipStream_t s) {
    const int fp32 = pyr.storage == VCT_STORE_FP32 ? 1 : 0;
    int src = 0;
    while (src + 1 < pyr.nlevels) {
        const LevelInfo &S = pyr.lv[src];
        const int R = S.size < 8 ? S.size : 8;
        if (R < 2) break;
        const int lg = R == 8 ? 3 : (R == 4 ? 2 : 1);
        int nout = pyr.nlevels - 1 - src;
        if (nout > lg) nout = lg;
        MipArgs a;
        a.pyr = pyr.pyr;
        a.src = S;
        for (int k = 0; k < 3; ++k) a.out[k] = pyr.lv[k < nout ? src + 1 + k : src + 1];
        a.nout = nout;
        a.fp32 = fp32;
        a.zero0 = src == 0 ? pyr.zero0 : nullptr;
        const int nreg = S.size / R;
        dim3 grid((unsigned)(nreg * nreg * nreg));
        if (pyr.aniso && S.nfaces == 6) hipLaunchKernelGGL((k_mip_chain<true, true>), grid, dim3(256), 0, s, a);
        else if (pyr.aniso) hipLaunchKernelGGL((k_mip_chain<true, false>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_mip_chain<false, false>), grid, dim3(256), 0, s, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        src += nout;
    }
    if (pyr.storage == VCT_STORE_FP8) {
        for (int l = 3; l < pyr.nlevels; ++l) {
            const LevelInfo &L = pyr.lv[l];
            hipLaunchKernelGGL(k_fp8_encode, dim3((unsigned)(L.nb * L.nb * L.nb * L.nfaces)), dim3(64), 0, s, pyr.pyr,
                               pyr.scales, L);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

}  // namespace vctk
