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

template <int MODE>
__device__ __forceinline__ float4 fetch(const LevelInfo &L, int face, int l, uint32_t a) {
    if (MODE == 1) return ((const float4 *)L.base[face])[a];
    if (MODE == 2 && l >= 3) {
        const uint32_t c = ((const uint32_t *)L.base[face])[a];
        const float sc = L.scale[face][a >> 6];
        return make_float4(fp82f(c & 0xff) * sc, fp82f((c >> 8) & 0xff) * sc, fp82f((c >> 16) & 0xff) * sc,
                           fp82f(c >> 24) * sc);
    }
    const uint2 v = ((const uint2 *)L.base[face])[a];
    return make_float4(h2f((uint16_t)(v.x & 0xffff)), h2f((uint16_t)(v.x >> 16)), h2f((uint16_t)(v.y & 0xffff)),
                       h2f((uint16_t)(v.y >> 16)));
}

struct Corner {
    int i0[3];
    float f[3];
    int s;
};

__device__ __forceinline__ Corner make_corner(int s, float px, float py, float pz) {
    Corner c;
    c.s = s;
    const float u[3] = {px * (float)s - 0.5f, py * (float)s - 0.5f, pz * (float)s - 0.5f};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float fl = __builtin_floorf(u[k]);
        c.i0[k] = (int)fl;
        c.f[k] = u[k] - fl;
    }
    return c;
}

// GL trilinear inside one level/face, corners in GL spec order.
template <int MODE>
__device__ __forceinline__ float4 linear(const LevelInfo &L, int face, int l, const Corner &c) {
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int s = c.s, nb = L.nb;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        const int x = c.i0[0] + dx, y = c.i0[1] + dy, z = c.i0[2] + dz;
        const float w = ((dx ? c.f[0] : 1.0f - c.f[0]) * (dy ? c.f[1] : 1.0f - c.f[1])) * (dz ? c.f[2] : 1.0f - c.f[2]);
        if ((unsigned)x < (unsigned)s && (unsigned)y < (unsigned)s && (unsigned)z < (unsigned)s) {
            const float4 t = fetch<MODE>(L, face, l, brick_addr(x, y, z, nb));
            acc.x += w * t.x;
            acc.y += w * t.y;
            acc.z += w * t.z;
            acc.w += w * t.w;
        }
    }
    return acc;
}

template <int MODE, bool ANISO>
__device__ __forceinline__ float4 level_sample(const LevelInfo *__restrict__ lv, int l, const float p[3],
                                               const float dir[3]) {
    const LevelInfo &L = lv[l];
    const Corner c = make_corner(L.size, p[0], p[1], p[2]);
    if (!ANISO || l == 0) return linear<MODE>(L, 0, l, c);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float w = dir[ax] * dir[ax];
        const int face = ax * 2 + (dir[ax] >= 0.0f ? 0 : 1);
        const float4 t = linear<MODE>(L, face, l, c);
        acc.x += w * t.x;
        acc.y += w * t.y;
        acc.z += w * t.z;
        acc.w += w * t.w;
    }
    return acc;
}

template <int MODE, bool ANISO>
__device__ __forceinline__ float4 sample_lod(const LevelInfo *__restrict__ lv, int dim, int nlevels, const float p[3],
                                             const float dir[3], float lambda) {
    const int q = nlevels - 1;
    if (lambda <= 0.0f) {  // magnification: NEAREST on level 0
        const int x = (int)__builtin_floorf(p[0] * (float)dim), y = (int)__builtin_floorf(p[1] * (float)dim),
                  z = (int)__builtin_floorf(p[2] * (float)dim);
        if ((unsigned)x >= (unsigned)dim || (unsigned)y >= (unsigned)dim || (unsigned)z >= (unsigned)dim)
            return make_float4(0.f, 0.f, 0.f, 0.f);
        const LevelInfo &L = lv[0];
        return fetch<MODE>(L, 0, 0, brick_addr(x, y, z, L.nb));
    }
    if (!(lambda < (float)q)) return level_sample<MODE, ANISO>(lv, q, p, dir);  // also catches NaN
    const float fl = __builtin_floorf(lambda);
    int d1 = (int)fl;
    d1 = d1 < 0 ? 0 : (d1 > q - 1 ? q - 1 : d1);
    const float fr = lambda - fl;
    const float4 t1 = level_sample<MODE, ANISO>(lv, d1, p, dir);
    const float4 t2 = level_sample<MODE, ANISO>(lv, d1 + 1, p, dir);
    return make_float4((1.0f - fr) * t1.x + fr * t2.x, (1.0f - fr) * t1.y + fr * t2.y,
                       (1.0f - fr) * t1.z + fr * t2.z, (1.0f - fr) * t1.w + fr * t2.w);
}

// traceCone (phong.frag:135-180)
template <int MODE, bool ANISO>
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
        const float4 s = sample_lod<MODE, ANISO>(lv, dim, nlevels, sp, dir, lod + lod_offset);
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
    const size_t pi = (size_t)py * t.width + px;
    const float4 *g = (const float4 *)&t.gbuf[pi];
    const float4 g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
    if (__float_as_uint(g2.w) == 0u) {
        t.out_ind[pi] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (t.enable_reflections) t.out_spec[pi] = make_float4(0.f, 0.f, 0.f, 0.f);
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
        const float4 r = trace_cone<MODE, ANISO>(lv, dim, nlevels, pos, N, d, t.d_steps, t.d_bias, t.d_tan_half,
                                                 t.d_h0, t.d_lod_offset);
        const float w = kConeW[i];
        i0 += w * r.x;
        i1 += w * r.y;
        i2 += w * r.z;
        i3 += w * r.w;
    }
    const float occ = 1.0f - clampf(i3, 0.0f, 1.0f);
    const float m = t.draw_occlusion ? occ : 1.0f;
    t.out_ind[pi] = make_float4(i0 * m, i1 * m, i2 * m, occ);
    if (t.enable_reflections) {
        float angle = t.s_angle;
        if (t.angle_from_roughness && t.has_roughness) angle = (g1.w * 3.1415982f) * 0.1f;  // PI of common.glsl:1
        const float I[3] = {gp[0] - t.eye[0], gp[1] - t.eye[1], gp[2] - t.eye[2]};
        const float dni = dot3(N, I);
        float R[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) R[k] = I[k] - (2.0f * dni) * N[k];
        const float4 r = trace_cone<MODE, ANISO>(lv, dim, nlevels, pos, N, R, t.s_steps, t.s_bias,
                                                 tanf(angle / 2.0f), t.s_h0, t.s_lod_offset);
        t.out_spec[pi] = r;
    }
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
