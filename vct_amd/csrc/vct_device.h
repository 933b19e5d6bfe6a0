// vct_device.h -- device-side helpers shared by the gfx950 kernels.
//
// Float arithmetic on the bit-exact passes (raster, voxelize, inject, mip) is
// written in one fixed evaluation order and compiled without FP contraction
// (each .hip file that includes this header for those passes sets
// `#pragma clang fp contract(off)`), so a pass's integer outputs and storage
// rounding are reproducible bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "vct_internal.h"

namespace vctk {

// ---- small vector math (fixed order, no fused multiply-add) ---------------
__device__ __forceinline__ float dot3(const float a[3], const float b[3]) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
// a / b correctly rounded to float (the IEEE division a / b), for finite
// floats a and b != 0 whose quotient is a normal float or zero, given rcp =
// RN64(1 / b).  q = RN64(a * rcp) is within (1 + 2^-53)^2 - 1 < 2^-52 of a / b,
// relatively.  A rounding boundary of float is a midpoint m = M 2^k, M odd
// with 25 bits; with |a| = A 2^i, |b| = B 2^j (A, B < 2^24), |a| - m |b| is a
// nonzero multiple of 2^min(i, k + j) (it is never zero: A would need the odd
// factor M), so ||a / b| - m| >= 2^-50 |a / b|.  q therefore rounds to float
// on the same side of every midpoint as a / b (rounding to nearest is
// symmetric in sign): RN32(q) = RN32(a / b).
__device__ __forceinline__ float quot_f32(float a, double rcp) { return (float)((double)a * rcp); }

// v / length(v), each component the IEEE quotient v[k] / l (quot_f32 with one
// FP64 reciprocal instead of three float divisions)
__device__ __forceinline__ void nrm3(const float v[3], float o[3]) {
    const float l = __builtin_sqrtf(dot3(v, v));
    const double r = 1.0 / (double)l;
    o[0] = quot_f32(v[0], r);
    o[1] = quot_f32(v[1], r);
    o[2] = quot_f32(v[2], r);
}
__device__ __forceinline__ void cross3(const float a[3], const float b[3], float o[3]) {
    float t0 = a[1] * b[2] - a[2] * b[1];
    float t1 = a[2] * b[0] - a[0] * b[2];
    float t2 = a[0] * b[1] - a[1] * b[0];
    o[0] = t0;
    o[1] = t1;
    o[2] = t2;
}
// column-major mat4 * vec4
__device__ __forceinline__ void mat4_vec4(const float *m, const float v[4], float o[4]) {
    float t[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float s = m[0 * 4 + r] * v[0];
        s = s + m[1 * 4 + r] * v[1];
        s = s + m[2 * 4 + r] * v[2];
        s = s + m[3 * 4 + r] * v[3];
        t[r] = s;
    }
    o[0] = t[0];
    o[1] = t[1];
    o[2] = t[2];
    o[3] = t[3];
}
__device__ __forceinline__ void mat3_vec3(const float *m, const float v[3], float o[3]) {
    float t[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) t[r] = (m[0 * 3 + r] * v[0] + m[1 * 3 + r] * v[1]) + m[2 * 3 + r] * v[2];
    o[0] = t[0];
    o[1] = t[1];
    o[2] = t[2];
}
__device__ __forceinline__ float interp3(const float l[3], float a0, float a1, float a2) {
    return (l[0] * a0 + l[1] * a1) + l[2] * a2;
}
__device__ __forceinline__ float clampf(float x, float a, float b) { return x < a ? a : (x > b ? b : x); }

// ---- warp texture --------------------------------------------------------------
// The 32^3 warp texture (warp.hip) as float4 texels, z-major: texture(warpmap,
// tc).xyz with LINEAR filtering and CLAMP_TO_EDGE (common.glsl:48-50,
// phong.frag:150-152, voxelize.frag:103-105).  Weights and sums in a fixed
// order (the oracle's or_warp_sample).
__device__ __forceinline__ void warp_sample(const float4 *__restrict__ wm, const float tc[3], float out[3]) {
    int i0[3], i1[3];
    float f[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float u = tc[k] * (float)kWarpDim - 0.5f;
        const float fl = __builtin_floorf(u);
        f[k] = u - fl;
        const int a = (int)fl, b = a + 1;
        i0[k] = a < 0 ? 0 : (a > kWarpDim - 1 ? kWarpDim - 1 : a);
        i1[k] = b < 0 ? 0 : (b > kWarpDim - 1 ? kWarpDim - 1 : b);
    }
    float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
        const float w = ((dx ? f[0] : 1.0f - f[0]) * (dy ? f[1] : 1.0f - f[1])) * (dz ? f[2] : 1.0f - f[2]);
        const int x = dx ? i1[0] : i0[0], y = dy ? i1[1] : i0[1], z = dz ? i1[2] : i0[2];
        const float4 t = wm[(z * kWarpDim + y) * kWarpDim + x];
        acc[0] = acc[0] + w * t.x;
        acc[1] = acc[1] + w * t.y;
        acc[2] = acc[2] + w * t.z;
    }
    out[0] = acc[0];
    out[1] = acc[1];
    out[2] = acc[2];
}

// ---- storage formats -------------------------------------------------------
__device__ __forceinline__ uint32_t pack_unorm4x8(float r, float g, float b, float a) {
    float c[4] = {r, g, b, a};
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float v = clampf(c[i], 0.0f, 1.0f);
        uint32_t q = (uint32_t)__builtin_rintf(v * 255.0f);
        o |= (q & 0xffu) << (8 * i);
    }
    return o;
}
// Byte i of v divided by 255, correctly rounded (what the IEEE division
// gives), in three operations instead of the division's ten: the product with
// the rounded reciprocal is one ulp off for 126 of the 256 bytes, and one
// correction step on the exact residual (two fmas) gives the correctly rounded
// quotient for every byte 0..255 (checked exhaustively in exact arithmetic).

__device__ __forceinline__ float unorm8(uint32_t v, int i) {
    const float x = (float)((v >> (8 * i)) & 0xffu);
    constexpr float R = 1.0f / 255.0f;  // the reciprocal rounded to float
    const float q = x * R;
    return fmaf(fmaf(-q, 255.0f, x), R, q);
}

__device__ __forceinline__ uint16_t f2h(float f) { return __half_as_ushort(__float2half_rn(f)); }
__device__ __forceinline__ float h2f(uint16_t h) { return __half2float(__ushort_as_half(h)); }
__device__ __forceinline__ float q16(float f) { return h2f(f2h(f)); }

// OCP FP8 E4M3: bias 7, max 448, 0x7f = NaN; round to nearest even, saturate.
__device__ __forceinline__ uint32_t f2fp8(float f) {
    if (f != f) return 0x7fu;
    uint32_t sign = (__float_as_uint(f) >> 24) & 0x80u;
    float a = __builtin_fabsf(f);
    if (a >= 448.0f) return sign | 0x7eu;
    if (a < 0.015625f) return sign | (uint32_t)__builtin_rintf(a * 512.0f);
    uint32_t bits = __float_as_uint(a);
    int e = (int)((bits >> 23) & 0xff) - 127;  // a = 1.m * 2^e, e >= -6
    float mant = __uint_as_float((bits & 0x7fffffu) | 0x3f800000u);  // [1,2)
    float q = __builtin_rintf((mant - 1.0f) * 8.0f);
    if (q >= 8.0f) {
        q = 0.0f;
        e += 1;
    }
    int code = ((e + 7) << 3) | (int)q;
    if (code > 0x7e) code = 0x7e;
    return sign | (uint32_t)code;
}
__device__ __forceinline__ float fp82f(uint32_t v) {
    uint32_t e = (v >> 3) & 15u, m = v & 7u;
    float mag;
    if (e == 0) mag = (float)m * 0.001953125f;  // m * 2^-9
    else mag = __uint_as_float(((e - 7 + 127) << 23) | (m << 20));
    return (v & 0x80u) ? -mag : mag;
}

// ---- voxel grid addressing ---------------------------------------------------
// The albedo and normal grids (dim^3 packed RGBA8) are stored in 4x4x4 bricks
// of 256 bytes, bricks and the voxels inside a brick x-fastest (nb = dim / 4):
// the voxels one fragment neighbourhood touches share a few cache lines.
__device__ __forceinline__ uint32_t grid_index(int x, int y, int z, int nb) {
    return ((uint32_t)(((z >> 2) * nb + (y >> 2)) * nb + (x >> 2)) << 6) | ((uint32_t)(z & 3) << 4) |
           ((uint32_t)(y & 3) << 2) | (uint32_t)(x & 3);
}

// ---- radiance pyramid addressing (layout in vct_internal.h) ----------------
__device__ __forceinline__ uint32_t elem_index(const LevelInfo &L, int x, int y, int z, int f) {
    return (((uint32_t)(x >> 2) * L.sx + (uint32_t)(x & 3)) + ((uint32_t)(y >> 2) * L.sy + (uint32_t)(y & 3) * 4u)) +
           (((uint32_t)(z >> 2) * L.sz + (uint32_t)(z & 3) * 16u) + (uint32_t)f * 64u);
}
__device__ __forceinline__ uint2 pack_h4(float4 c) {
    return make_uint2((uint32_t)f2h(c.x) | ((uint32_t)f2h(c.y) << 16), (uint32_t)f2h(c.z) | ((uint32_t)f2h(c.w) << 16));
}
__device__ __forceinline__ float4 unpack_h4(uint32_t a, uint32_t b) {
    return make_float4(h2f((uint16_t)(a & 0xffffu)), h2f((uint16_t)(a >> 16)), h2f((uint16_t)(b & 0xffffu)),
                       h2f((uint16_t)(b >> 16)));
}
__device__ __forceinline__ float4 decode_fp8x4(uint32_t c, float sc) {
    return make_float4(fp82f(c & 0xffu) * sc, fp82f((c >> 8) & 0xffu) * sc, fp82f((c >> 16) & 0xffu) * sc,
                       fp82f(c >> 24) * sc);
}

// Texel (x,y,z) of face f in a level's stage (fp32, or fp16 pairs).
__device__ __forceinline__ float4 load_stage_texel(const char *pyr, const LevelInfo &L, int x, int y, int z, int f) {
    const char *p = pyr + L.stage + (size_t)elem_index(L, x, y, z, f) * 16;
    if (L.fmt == 1) return *(const float4 *)p;
    const uint2 v = *(const uint2 *)p;
    return unpack_h4(v.x, v.y);
}
// Writes texel (x,y,z) of face f into a level's stage: the fp32 element, or
// both fp16 halves that hold it (the first half of its own element and the
// second half of the element of x - 1, in the border brick for x = 0; its
// element index wraps below 0, so the byte offset is formed in 32 bits).
__device__ __forceinline__ void store_stage_texel(char *pyr, const LevelInfo &L, int x, int y, int z, int f,
                                                  float4 c) {
    char *p = pyr + L.stage + (size_t)elem_index(L, x, y, z, f) * 16;
    if (L.fmt == 1) {
        *(float4 *)p = c;
        return;
    }
    const uint2 h = pack_h4(c);
    *(uint2 *)p = h;
    *(uint2 *)(pyr + (uint32_t)(L.stage + elem_index(L, x - 1, y, z, f) * 16u) + 8) = h;
}
// Texel (x,y,z) of face f in a level's storage, decoded to float.
__device__ __forceinline__ float4 load_texel(const char *pyr, const float *scales, const LevelInfo &L, int x, int y,
                                             int z, int f) {
    const uint32_t e = elem_index(L, x, y, z, f);
    if (L.fmt == 2) return decode_fp8x4(*(const uint32_t *)(pyr + L.base + (size_t)e * 4), scales[L.scale + (e >> 6)]);
    const char *p = pyr + L.base + (size_t)e * 16;
    if (L.fmt == 1) return *(const float4 *)p;
    const uint2 v = *(const uint2 *)p;
    return unpack_h4(v.x, v.y);
}

// GL LINEAR inside face 0 of one level (border 0), FP32 weights and sums.
// p in [0,1]^3 texture coordinates; s = the level's edge.
__device__ __forceinline__ float4 level_linear(const char *pyr, const float *scales, const LevelInfo &L, const float p[3]) {
    const int s = L.size;
    int i0[3];
    float f[3];
    for (int k = 0; k < 3; ++k) {
        const float u = p[k] * (float)s - 0.5f;
        const float fl = __builtin_floorf(u);
        i0[k] = (int)fl;
        f[k] = u - fl;
    }
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (L.fmt == 0) {  // fp16 pairs: one load per (y, z) corner pair
        const bool xin = i0[0] >= 0 && i0[0] < s;  // i0 = s - 1 reads the zero border half
        const bool xneg = i0[0] == -1;
        if (!xin && !xneg) return acc;
        const int cx = xneg ? 0 : i0[0];
        const float wlo = xneg ? f[0] : 1.0f - f[0], whi = xneg ? 0.0f : f[0];
        for (int c = 0; c < 4; ++c) {
            const int dy = c & 1, dz = c >> 1;
            const int y = i0[1] + dy, z = i0[2] + dz;
            if ((unsigned)y >= (unsigned)s || (unsigned)z >= (unsigned)s) continue;
            const float wyz = (dy ? f[1] : 1.0f - f[1]) * (dz ? f[2] : 1.0f - f[2]);
            const uint4 v = *(const uint4 *)(pyr + L.base + (size_t)elem_index(L, cx, y, z, 0) * 16);
            const float4 lo = unpack_h4(v.x, v.y), hi = unpack_h4(v.z, v.w);
            const float a = wlo * wyz, b = whi * wyz;
            acc.x += a * lo.x + b * hi.x;
            acc.y += a * lo.y + b * hi.y;
            acc.z += a * lo.z + b * hi.z;
            acc.w += a * lo.w + b * hi.w;
        }
        return acc;
    }
    for (int c = 0; c < 8; ++c) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
        const int x = i0[0] + dx, y = i0[1] + dy, z = i0[2] + dz;
        if ((unsigned)x >= (unsigned)s || (unsigned)y >= (unsigned)s || (unsigned)z >= (unsigned)s) continue;
        const float w = ((dx ? f[0] : 1.0f - f[0]) * (dy ? f[1] : 1.0f - f[1])) * (dz ? f[2] : 1.0f - f[2]);
        const float4 t = load_texel(pyr, scales, L, x, y, z, 0);
        acc.x += w * t.x;
        acc.y += w * t.y;
        acc.z += w * t.z;
        acc.w += w * t.w;
    }
    return acc;
}

// GL LINEAR sample of an fp32 depth texture, CLAMP_TO_BORDER with border 1.0
// (src/Application.cpp:46-53), with a textureOffset texel offset.
__device__ __forceinline__ float shadow_bilinear(const float *__restrict__ sm, int N, float s, float t, int ox,
                                                 int oy) {
    float u = s * (float)N;
    u = u + (float)ox;
    u = u - 0.5f;
    float v = t * (float)N;
    v = v + (float)oy;
    v = v - 0.5f;
    const float fu = __builtin_floorf(u), fv = __builtin_floorf(v);
    const int i0 = (int)fu, j0 = (int)fv;
    const float a = u - fu, b = v - fv;
    auto tex = [&](int i, int j) -> float {
        return (i >= 0 && i < N && j >= 0 && j < N) ? sm[(size_t)j * N + i] : 1.0f;
    };
    const float t00 = tex(i0, j0), t10 = tex(i0 + 1, j0), t01 = tex(i0, j0 + 1), t11 = tex(i0 + 1, j0 + 1);
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    return ((w00 * t00 + w10 * t10) + w01 * t01) + w11 * t11;
}

// calcShadowFactor (shaders/voxelize.frag:159-180, phong.frag:183-207): the
// fraction of the five PCF taps (offsets (0,0), (1,0), (0,1), (-1,0), (0,-1))
// that occlude a light-space position; 0 beyond the far plane.  In two halves
// so that a caller can have the taps of several lights in flight at once:
// shadow_taps fetches the five footprints, shadow_count compares them (the
// same operations, in the same order, as one shadow_bilinear_quad per tap).
struct ShadowTaps {
    float4 q[5];  // the footprints of the five taps (1.0 outside the map)
    float s, t, frag;
    bool far;  // beyond the far plane: factor 0, nothing fetched
};
__device__ __forceinline__ void shadow_tap_index(int N, float s, float t, int ox, int oy, int *i0, int *j0) {
    float u = s * (float)N;
    u = u + (float)ox;
    u = u - 0.5f;
    float v = t * (float)N;
    v = v + (float)oy;
    v = v - 0.5f;
    *i0 = (int)__builtin_floorf(u);
    *j0 = (int)__builtin_floorf(v);
}
__device__ __forceinline__ void shadow_taps(const float lsp[4], const float4 *sq, int N, ShadowTaps &st) {
    float sh[3];
    if (__ballot(lsp[3] != 1.0f) == 0) {  // orthographic light space: w = 1, and x / 1 is x
        for (int k = 0; k < 3; ++k) sh[k] = (lsp[k] + 1.0f) * 0.5f;
    } else {
        const double rw = 1.0 / (double)lsp[3];  // the three IEEE quotients by w (quot_f32)
        for (int k = 0; k < 3; ++k) sh[k] = (quot_f32(lsp[k], rw) + 1.0f) * 0.5f;
    }
    st.s = sh[0];
    st.t = sh[1];
    st.frag = sh[2] - 0.01f;
    st.far = st.frag > 1.0f;
    const int ox[5] = {0, 1, 0, -1, 0}, oy[5] = {0, 0, 1, 0, -1};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        int i0, j0;
        shadow_tap_index(N, st.s, st.t, ox[k], oy[k], &i0, &j0);
        st.q[k] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        if (!st.far && i0 >= -1 && i0 <= N - 1 && j0 >= -1 && j0 <= N - 1)
            st.q[k] = sq[(size_t)(j0 + 1) * (N + 1) + (i0 + 1)];
    }
}
__device__ __forceinline__ float shadow_count(const ShadowTaps &st, int N) {
    if (st.far) return 0.0f;
    const int ox[5] = {0, 1, 0, -1, 0}, oy[5] = {0, 0, 1, 0, -1};
    float f = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        float u = st.s * (float)N;
        u = u + (float)ox[k];
        u = u - 0.5f;
        float v = st.t * (float)N;
        v = v + (float)oy[k];
        v = v - 0.5f;
        const float a = u - __builtin_floorf(u), b = v - __builtin_floorf(v);
        const float4 T = st.q[k];
        const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
        if (st.frag > ((w00 * T.x + w10 * T.y) + w01 * T.z) + w11 * T.w) f += 1.0f;
    }
    return f * 0.2f;  // the same float as f / 5 for each of f = 0, 1, ..., 5
}
__device__ __forceinline__ float calc_shadow_factor(const float lsp[4], const float4 *sq, int N) {
    ShadowTaps st;
    shadow_taps(lsp, sq, N, st);
    return shadow_count(st, N);
}

// ---- wave helpers ----------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Reduces (sum a, max b, sum c, sum d) over a 256-thread workgroup and writes
// them to partials[blockIdx.x * 4 ..].  Every thread of the block must call it.
__device__ __forceinline__ void block_partial(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t *partials) {
    __shared__ uint32_t red[4][4];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        const uint32_t ob = __shfl_xor(b, off, 64);
        b = ob > b ? ob : b;
        c += __shfl_xor(c, off, 64);
        d += __shfl_xor(d, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w][0] = a;
        red[w][1] = b;
        red[w][2] = c;
        red[w][3] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        uint32_t s0 = 0, m1 = 0, s2 = 0, s3 = 0;
        for (int i = 0; i < nw; ++i) {
            s0 += red[i][0];
            m1 = red[i][1] > m1 ? red[i][1] : m1;
            s2 += red[i][2];
            s3 += red[i][3];
        }
        uint32_t *o = partials + (size_t)blockIdx.x * 4;
        o[0] = s0;
        o[1] = m1;
        o[2] = s2;
        o[3] = s3;
    }
}

}  // namespace vctk
