// texture.h -- the diffuse-map sampler of the voxelizer and the G-buffer.
//
// GL 4.5 core 8.14 with the sampler state the reference sets on every diffuse
// map (src/ResourceLoader.h:79-82, src/Graphics/GLHelper.cpp:180-183):
//   lambda = log2(rho), rho^2 = max((du/dx)^2 + (dv/dx)^2, (du/dy)^2 + (dv/dy)^2)
//     with u, v = s, t times level 0's size (8.14.1; the caller passes rho^2);
//   lambda <= 0 (rho^2 <= 1, and a zero footprint): magnification, NEAREST on
//     level 0: i = floor(u) (8.14.2);
//   lambda > 0: LINEAR_MIPMAP_NEAREST, level d = 0 for lambda <= 1/2, else
//     ceil(lambda + 1/2) - 1, clamped to q = floor(log2(max size)) (8.14.3);
//     lambda in (k - 1/2, k + 1/2] is rho^2 in (2^(2k-1), 2^(2k+1)], so d is
//     found by comparing rho^2 with powers of two -- no logarithm, the same
//     level wherever the comparison runs;
//   LINEAR on level d: i0 = floor(u - 1/2), j0 = floor(v - 1/2), a and b their
//     fractions, tau = (1-a)(1-b) t00 + a(1-b) t10 + (1-a)b t01 + ab t11;
//   REPEAT: an index i is taken modulo the level's size;
//   texels are unsigned normalized (byte / 255);
//   a mipmap-incomplete texture (fewer levels than q + 1) or an index outside
//     the table samples (0, 0, 0, 1).
// The arithmetic is FP32 in this order, without contraction, in every caller.
#pragma once
#include "vct_device.h"

namespace vctk {

__device__ __forceinline__ int tex_wrap(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

// floor(x) as an int, x clamped to +-2^30 first (coordinates that large are
// whole numbers in FP32; the clamp keeps the conversion defined)
__device__ __forceinline__ int tex_floor(float x) {
    const float f = __builtin_floorf(x);
    return (int)fminf(fmaxf(f, -1073741824.0f), 1073741824.0f);
}

__device__ __forceinline__ float4 tex_unorm(uint32_t v) {
    return make_float4(unorm8(v, 0), unorm8(v, 1), unorm8(v, 2), unorm8(v, 3));
}

__device__ inline float4 tex_sample(const DevTex &tx, int id, float s, float t, float rho2) {
    if (id < 0 || id >= tx.count) return make_float4(0.f, 0.f, 0.f, 1.f);
    const int4 inf = tx.info[id];
    if (!(inf.w & 1)) return make_float4(0.f, 0.f, 0.f, 1.f);
    const int q = (inf.w >> 8) & 0xff;
    if (!(rho2 > 1.0f)) {  // magnification: NEAREST on level 0
        const int4 L = tx.levels[inf.x];
        const int i = tex_wrap(tex_floor(s * (float)L.y), L.y), j = tex_wrap(tex_floor(t * (float)L.z), L.z);
        return tex_unorm(tx.texels[(size_t)(uint32_t)L.x + (size_t)j * L.y + i]);
    }
    int d = 0;
    for (int k = 1; k <= q; ++k)
        if (rho2 > __builtin_ldexpf(1.0f, 2 * k - 1)) d = k;
    const int4 L = tx.levels[inf.x + d];
    const uint32_t *T = tx.texels + (uint32_t)L.x;
    const float u = s * (float)L.y - 0.5f, v = t * (float)L.z - 0.5f;
    const float fu = __builtin_floorf(u), fv = __builtin_floorf(v);
    const float a = u - fu, b = v - fv;
    const int i0 = tex_floor(u), j0 = tex_floor(v);
    const int ia = tex_wrap(i0, L.y), ib = tex_wrap(i0 + 1, L.y);
    const int ja = tex_wrap(j0, L.z), jb = tex_wrap(j0 + 1, L.z);
    const float4 t00 = tex_unorm(T[(size_t)ja * L.y + ia]), t10 = tex_unorm(T[(size_t)ja * L.y + ib]);
    const float4 t01 = tex_unorm(T[(size_t)jb * L.y + ia]), t11 = tex_unorm(T[(size_t)jb * L.y + ib]);
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    float4 r;
    r.x = ((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x;
    r.y = ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y;
    r.z = ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z;
    r.w = ((w00 * t00.w + w10 * t10.w) + w01 * t01.w) + w11 * t11.w;
    return r;
}

// rho^2 of an affine (s, t) over a triangle in window pixels (orthographic
// projection: the voxelizer's dominant-axis views): the plane through the
// three corners, its s and t derivatives scaled by level 0's size.
__device__ __forceinline__ float tex_rho2_affine(const float win[3][2], const float st[6], float W, float H) {
    const float e1x = win[1][0] - win[0][0], e1y = win[1][1] - win[0][1];
    const float e2x = win[2][0] - win[0][0], e2y = win[2][1] - win[0][1];
    const float den = e1x * e2y - e2x * e1y;
    const float du1 = (st[2] - st[0]) * W, du2 = (st[4] - st[0]) * W;
    const float dv1 = (st[3] - st[1]) * H, dv2 = (st[5] - st[1]) * H;
    const float dudx = (du1 * e2y - du2 * e1y) / den, dvdx = (dv1 * e2y - dv2 * e1y) / den;
    const float dudy = (du2 * e1x - du1 * e2x) / den, dvdy = (dv2 * e1x - dv1 * e2x) / den;
    return fmaxf(dudx * dudx + dvdx * dvdx, dudy * dudy + dvdy * dvdy);
}

}  // namespace vctk
