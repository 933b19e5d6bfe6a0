// trace.hip -- the cone-traced indirect light of phong.frag: traceCone
// (shaders/phong.frag:135-180), the 6-cone diffuse gather with its occlusion
// term and the specular cone (:455-512).
//
// Layout of the work: a workgroup is a 16x16 pixel tile made of four 8x8
// wavefront tiles, so the 64 lanes of a wavefront march neighbouring cones
// through the same bricks.  Workgroup ids are remapped so each of the eight
// XCDs (which receive workgroups round-robin) owns one contiguous band of
// tiles and keeps that band's bricks in its own L2.
//
// Sampling restates the reference sampler (src/Application.h:144-146,
// Application.cpp:1094-1098): LINEAR_MIPMAP_LINEAR with the lambda clamp to
// the last level, trilinear inside a level, border (0,0,0,0) outside.
// Accumulation is FP32 whatever the storage format.
#include "vct_device.h"

namespace vctk {

// ---- trilinear sampling ------------------------------------------------------
// A level of edge s is stored as 4^3 bricks (nb = ceil(s/4) per axis), so a
// texel address is the sum of three per-axis terms:
//   addr(x,y,z) = X(x) + Y(y) + Z(z),  X(x) = (x>>2)*64 + (x&3),
//   Y(y) = (y>>2)*nb*64 + (y&3)*4,     Z(z) = (z>>2)*nb*nb*64 + (z&3)*16.
// One corner setup per level gives the 8 addresses and weights; corners
// outside the level (the border) are clamped to a valid texel and get weight
// 0, so the 8 fetches are unconditional.
struct Corners {
    uint32_t a[8];
    float w[8];
};

__device__ __forceinline__ void axis_terms(float p, int s, uint32_t mul, uint32_t sub, uint32_t t[2], float w[2]) {
    const float u = p * (float)s - 0.5f;
    const float fl = __builtin_floorf(u);
    const int i0 = (int)fl;
    const float f = u - fl;
    const int i1 = i0 + 1;
    w[0] = (i0 >= 0) ? 1.0f - f : 0.0f;  // i0 <= s - 1 always (p <= 1)
    w[1] = (i1 < s) ? f : 0.0f;          // i1 >= 0 always (p >= 0)
    const int c0 = i0 < 0 ? 0 : i0, c1 = i1 > s - 1 ? s - 1 : i1;
    t[0] = (uint32_t)(c0 >> 2) * mul + (uint32_t)(c0 & 3) * sub;
    t[1] = (uint32_t)(c1 >> 2) * mul + (uint32_t)(c1 & 3) * sub;
}

__device__ __forceinline__ void make_corners(int s, const float p[3], float level_w, Corners &c) {
    const uint32_t nb = (uint32_t)((s + 3) >> 2);
    uint32_t tx[2], ty[2], tz[2];
    float wx[2], wy[2], wz[2];
    axis_terms(p[0], s, 64u, 1u, tx, wx);
    axis_terms(p[1], s, nb * 64u, 4u, ty, wy);
    axis_terms(p[2], s, nb * nb * 64u, 16u, tz, wz);
    wz[0] *= level_w;
    wz[1] *= level_w;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        c.a[k] = tx[dx] + ty[dy] + tz[dz];
        c.w[k] = (wx[dx] * wy[dy]) * wz[dz];
    }
}

__device__ __forceinline__ void acc_half4(float4 &acc, uint2 v, float w) {
    acc.x = fmaf(__half2float(__ushort_as_half((uint16_t)(v.x & 0xffffu))), w, acc.x);
    acc.y = fmaf(__half2float(__ushort_as_half((uint16_t)(v.x >> 16))), w, acc.y);
    acc.z = fmaf(__half2float(__ushort_as_half((uint16_t)(v.y & 0xffffu))), w, acc.z);
    acc.w = fmaf(__half2float(__ushort_as_half((uint16_t)(v.y >> 16))), w, acc.w);
}

// acc += scale * sum_k c.w[k] * texel(face base, c.a[k])
template <int MODE>
__device__ __forceinline__ void gather(float4 &acc, const char *base, const float *scales, int l, const Corners &c,
                                       float scale) {
    if (MODE == 1) {
        const float4 *b = (const float4 *)base;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float4 t = b[c.a[k]];
            const float w = c.w[k] * scale;
            acc.x = fmaf(t.x, w, acc.x);
            acc.y = fmaf(t.y, w, acc.y);
            acc.z = fmaf(t.z, w, acc.z);
            acc.w = fmaf(t.w, w, acc.w);
        }
    } else if (MODE == 2 && l >= 3) {
        const uint32_t *b = (const uint32_t *)base;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t t = b[c.a[k]];
            const float w = (c.w[k] * scale) * scales[c.a[k] >> 6];
            acc.x = fmaf(fp82f(t & 0xffu), w, acc.x);
            acc.y = fmaf(fp82f((t >> 8) & 0xffu), w, acc.y);
            acc.z = fmaf(fp82f((t >> 16) & 0xffu), w, acc.z);
            acc.w = fmaf(fp82f(t >> 24), w, acc.w);
        }
    } else {
        const uint2 *b = (const uint2 *)base;
        uint2 t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) t[k] = b[c.a[k]];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc_half4(acc, t[k], c.w[k] * scale);
    }
}

// One level (weight level_w): isotropic, or the dir^2-weighted mix of the three
// faces that face the cone (anisotropic levels >= 1).
template <int MODE, bool ANISO>
__device__ __forceinline__ void level_accumulate(float4 &acc, const LevelInfo *__restrict__ lv, int dim, int l,
                                                 const float p[3], const float dir[3], float level_w) {
    const int s = (dim >> l) > 0 ? (dim >> l) : 1;
    Corners c;
    make_corners(s, p, level_w, c);
    const LevelInfo &L = lv[l];
    if (!ANISO || l == 0) {
        gather<MODE>(acc, L.base[0], L.scale[0], l, c, 1.0f);
        return;
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const bool pos = dir[ax] >= 0.0f;
        const char *base = pos ? L.base[ax * 2] : L.base[ax * 2 + 1];
        const float *sc = pos ? L.scale[ax * 2] : L.scale[ax * 2 + 1];
        gather<MODE>(acc, base, sc, l, c, dir[ax] * dir[ax]);
    }
}

// textureLod with LINEAR_MIPMAP_LINEAR and the lambda clamp to [0, q].
// ULOD: lambda is the same on every active lane, so the level choice and the
// per-level state are wavefront-uniform (scalar registers and loads).
template <int MODE, bool ANISO, bool ULOD>
__device__ __forceinline__ float4 sample_lod(const LevelInfo *__restrict__ lv, int dim, int nlevels, const float p[3],
                                             const float dir[3], float lambda) {
    if (ULOD) lambda = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(lambda)));
    const int q = nlevels - 1;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (lambda <= 0.0f) {  // magnification: NEAREST on level 0
        const int x = (int)__builtin_floorf(p[0] * (float)dim), y = (int)__builtin_floorf(p[1] * (float)dim),
                  z = (int)__builtin_floorf(p[2] * (float)dim);
        if ((unsigned)x >= (unsigned)dim || (unsigned)y >= (unsigned)dim || (unsigned)z >= (unsigned)dim) return acc;
        Corners c;
        const int nb = (dim + 3) >> 2;
        for (int k = 0; k < 8; ++k) {
            c.a[k] = brick_addr(x, y, z, nb);
            c.w[k] = k == 0 ? 1.0f : 0.0f;
        }
        gather<MODE>(acc, lv[0].base[0], lv[0].scale[0], 0, c, 1.0f);
        return acc;
    }
    if (!(lambda < (float)q)) {  // lambda >= q (or NaN): the last level only
        level_accumulate<MODE, ANISO>(acc, lv, dim, q, p, dir, 1.0f);
        return acc;
    }
    const float fl = __builtin_floorf(lambda);
    int d1 = (int)fl;
    d1 = d1 < 0 ? 0 : (d1 > q - 1 ? q - 1 : d1);
    const float fr = lambda - fl;
    level_accumulate<MODE, ANISO>(acc, lv, dim, d1, p, dir, 1.0f - fr);
    level_accumulate<MODE, ANISO>(acc, lv, dim, d1 + 1, p, dir, fr);
    return acc;
}

// traceCone (phong.frag:135-180).  ULOD: every lane of the wavefront traces a
// cone of the same aperture and initial height, so the step schedule h_i (and
// with it the sampled level) depends only on the step index.
template <int MODE, bool ANISO, bool ULOD>
__device__ float4 trace_cone(const LevelInfo *__restrict__ lv, int dim, int nlevels, const float pos[3],
                             const float nrm[3], const float dir_in[3], int steps, float bias, float tan_half,
                             float h0, float lod_offset) {
    float dir[3];
    nrm3(dir_in, dir);
    float c0 = 0.f, c1 = 0.f, c2 = 0.f, a = 0.f;
    const float scale = 1.0f / (float)dim;
    float start[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) start[k] = pos[k] + (bias * nrm[k]) * scale;
    float h = h0;
    for (int i = 0; i < steps && a < 0.95f; ++i) {
        const float r = h * tan_half;
        const float lod = log2f(fmaxf(1.0f, 2.0f * r));
        float sp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) sp[k] = start[k] + (h * dir[k]) * scale;
        if (sp[0] < 0.f || sp[0] > 1.f || sp[1] < 0.f || sp[1] > 1.f || sp[2] < 0.f || sp[2] > 1.f) break;
        const float4 s = sample_lod<MODE, ANISO, ULOD>(lv, dim, nlevels, sp, dir, lod + lod_offset);
        const float aa = 1.0f - a;
        c0 += s.x * aa;
        c1 += s.y * aa;
        c2 += s.z * aa;
        a += aa * s.w;
        h += r;
    }
    return make_float4(c0, c1, c2, a);
}

__constant__ float kConeDirs[6][3] = {{0.0f, 1.0f, 0.0f},
                                      {0.0f, 0.5f, 0.866025f},
                                      {0.823639f, 0.5f, 0.267617f},
                                      {0.509037f, 0.5f, -0.700629f},
                                      {-0.5909037f, 0.5f, -0.700629f},  // sic: phong.frag:475
                                      {-0.823639f, 0.5f, 0.267617f}};
__constant__ float kConeW[6] = {0.25f, 0.15f, 0.15f, 0.15f, 0.15f, 0.15f};

// One pixel of the phong.frag gather: 6 diffuse cones (+ the specular cone).
template <int MODE, bool ANISO>
__device__ __forceinline__ void trace_pixel(const TraceLaunch &t, const LevelInfo *__restrict__ lv, int dim,
                                            int nlevels, int px, int py, bool reflections) {
    const size_t pi = (size_t)py * t.width + px;
    const float4 *g = (const float4 *)&t.gbuf[pi];
    const float4 g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
    if (__float_as_uint(g2.w) == 0u) {
        t.out_ind[pi] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (reflections) t.out_spec[pi] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    float pos[3];
    const float gp[3] = {g0.x, g0.y, g0.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) pos[k] = ((gp[k] - t.ext.center[k]) - t.ext.vmin[k]) / (t.ext.vmax[k] - t.ext.vmin[k]);
    const float nraw[3] = {g1.x, g1.y, g1.z};
    const float T[3] = {g2.x, g2.y, g2.z};
    const float B[3] = {g3.x, g3.y, g3.z};
    float N[3];
    nrm3(nraw, N);
    float i0 = 0.f, i1 = 0.f, i2 = 0.f, i3 = 0.f;
    for (int i = 0; i < 6; ++i) {
        const float cx = kConeDirs[i][0], cy = kConeDirs[i][1], cz = kConeDirs[i][2];
        float d[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) d[k] = (T[k] * cx + B[k] * cy) + nraw[k] * cz;
        const float4 r = trace_cone<MODE, ANISO, true>(lv, dim, nlevels, pos, N, d, t.d_steps, t.d_bias,
                                                       t.d_tan_half, t.d_h0, t.d_lod_offset);
        const float w = kConeW[i];
        i0 += w * r.x;
        i1 += w * r.y;
        i2 += w * r.z;
        i3 += w * r.w;
    }
    const float occ = 1.0f - clampf(i3, 0.0f, 1.0f);
    const float m = t.draw_occlusion ? occ : 1.0f;
    t.out_ind[pi] = make_float4(i0 * m, i1 * m, i2 * m, occ);
    if (reflections) {
        float angle = t.s_angle;
        if (t.angle_from_roughness && t.has_roughness) angle = (g1.w * 3.1415982f) * 0.1f;  // PI of common.glsl:1
        const float I[3] = {gp[0] - t.eye[0], gp[1] - t.eye[1], gp[2] - t.eye[2]};
        const float dni = dot3(N, I);
        float R[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) R[k] = I[k] - (2.0f * dni) * N[k];
        const float4 r = trace_cone<MODE, ANISO, false>(lv, dim, nlevels, pos, N, R, t.s_steps, t.s_bias,
                                                        tanf(angle / 2.0f), t.s_h0, t.s_lod_offset);
        t.out_spec[pi] = r;
    }
}

template <int MODE, bool ANISO>
__global__ __launch_bounds__(256) void k_trace(TraceLaunch t, const LevelInfo *__restrict__ lv, int dim, int nlevels,
                                               int tiles_x, int tiles, int per_xcd) {
    // XCD-aware tile order: workgroup b runs on XCD b % 8; give XCD k the
    // contiguous tile band [k * per_xcd, (k + 1) * per_xcd).
    const int b = blockIdx.x;
    const int tile = (b & 7) * per_xcd + (b >> 3);
    if (tile >= tiles) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = (tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = (tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    if (px >= t.width || py >= t.height) return;
    trace_pixel<MODE, ANISO>(t, lv, dim, nlevels, px, py, t.enable_reflections != 0);
}

// Diffuse gather of listed 8x8 tiles (the learned gather's fallback): one
// wavefront per tile.
template <int MODE, bool ANISO>
__global__ __launch_bounds__(64) void k_trace_tiles(TraceLaunch t, const LevelInfo *__restrict__ lv, int dim,
                                                    int nlevels, const uint32_t *__restrict__ list,
                                                    const uint32_t *__restrict__ count) {
    const uint32_t n = *count;
    const int tiles_x = (t.width + 7) / 8;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int tile = (int)list[i];
        const int px = (tile % tiles_x) * 8 + (threadIdx.x & 7), py = (tile / tiles_x) * 8 + (threadIdx.x >> 3);
        if (px < t.width && py < t.height) trace_pixel<MODE, ANISO>(t, lv, dim, nlevels, px, py, false);
    }
}

hipError_t launch_trace_tiles(const PyrDesc &pyr, const TraceLaunch &t, const uint32_t *list, const uint32_t *count,
                              int max_tiles, hipStream_t s) {
    const int mode = pyr.storage == VCT_STORE_FP32 ? 1 : (pyr.storage == VCT_STORE_FP8 ? 2 : 0);
    dim3 grid((unsigned)(max_tiles < 8192 ? (max_tiles > 0 ? max_tiles : 1) : 8192));
#define VCT_TILES_LAUNCH(M, A) \
    hipLaunchKernelGGL((k_trace_tiles<M, A>), grid, dim3(64), 0, s, t, pyr.d_lv, pyr.dim, pyr.nlevels, list, count)
    if (pyr.aniso) {
        if (mode == 0) VCT_TILES_LAUNCH(0, true);
        else if (mode == 1) VCT_TILES_LAUNCH(1, true);
        else VCT_TILES_LAUNCH(2, true);
    } else {
        if (mode == 0) VCT_TILES_LAUNCH(0, false);
        else if (mode == 1) VCT_TILES_LAUNCH(1, false);
        else VCT_TILES_LAUNCH(2, false);
    }
#undef VCT_TILES_LAUNCH
    return hipGetLastError();
}

hipError_t launch_trace(const PyrDesc &pyr, const TraceLaunch &t, hipStream_t s) {
    const int tiles_x = (t.width + 15) / 16, tiles_y = (t.height + 15) / 16;
    const int tiles = tiles_x * tiles_y;
    const int per_xcd = (tiles + 7) / 8;
    dim3 grid((unsigned)(per_xcd * 8));
    const int mode = pyr.storage == VCT_STORE_FP32 ? 1 : (pyr.storage == VCT_STORE_FP8 ? 2 : 0);
#define VCT_TRACE_LAUNCH(M, A)                                                                                      \
    hipLaunchKernelGGL((k_trace<M, A>), grid, dim3(256), 0, s, t, pyr.d_lv, pyr.dim, pyr.nlevels, tiles_x, tiles, \
                       per_xcd)
    if (pyr.aniso) {
        if (mode == 0) VCT_TRACE_LAUNCH(0, true);
        else if (mode == 1) VCT_TRACE_LAUNCH(1, true);
        else VCT_TRACE_LAUNCH(2, true);
    } else {
        if (mode == 0) VCT_TRACE_LAUNCH(0, false);
        else if (mode == 1) VCT_TRACE_LAUNCH(1, false);
        else VCT_TRACE_LAUNCH(2, false);
    }
#undef VCT_TRACE_LAUNCH
    return hipGetLastError();
}

}  // namespace vctk
