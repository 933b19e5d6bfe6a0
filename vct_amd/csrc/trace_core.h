// trace_core.h -- the cone tracer's device code: the pyramid sampler,
// traceCone (shaders/phong.frag:135-180) and the per-pixel gather of
// phong.frag:455-512.  Shared by the trace kernels (trace.hip) and the learned
// gather's per-tile fallback (learned.hip).
//
// Sampling restates the reference sampler (src/Application.h:144-146,
// Application.cpp:1094-1098): LINEAR_MIPMAP_LINEAR with the lambda clamp to
// the last level, trilinear inside a level, border (0,0,0,0) outside.
// Accumulation is FP32 whatever the storage format.
#pragma once

#include "vct_device.h"

namespace vctk {

// ---- trilinear sampling ------------------------------------------------------
// Every fetch is a buffer load through one descriptor that spans the whole
// pyramid (kept in scalar registers), with a 32-bit byte offset per lane.  An
// element index is a sum of per-axis terms (vct_internal.h):
//   X(x) = (x>>2)*sx + (x&3),  Y(y) = (y>>2)*sy + (y&3)*4,
//   Z(z) = (z>>2)*sz + (z&3)*16, plus 64 per face.
// fp16 levels read corners outside the level (the border) from the zero
// border bricks around every level; the one-texel-per-element formats clamp
// them to a valid texel with weight 0.  Either way all fetches are
// unconditional.  fp16 levels store texel pairs
// (x, x+1) per element: a trilinear sample is four 16-byte loads, one per
// (y, z) corner pair, instead of eight 8-byte loads -- the vector memory
// pipeline processes a wavefront's load in the same number of cycles for 8 or
// 16 bytes per lane, so this halves the trace's dominant cost.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct PyrArgs {
    const char *pyr;
    uint32_t bytes;
    const float *scales;
    const LevelInfo *lv;
    int dim, nlevels;
    int nested;  // TraceLaunch::nested
};

__device__ __forceinline__ uint4 bload16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t bload4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
}
// Scalar load of 16 bytes at a wave-uniform offset: through the constant
// address space the load becomes s_load_dwordx4 into SGPRs (scalar data cache,
// not the vector memory pipeline).  The pyramid is not written while a trace
// runs.  The offset is one the vector path would load (clamped texel indices),
// so it is inside the allocation.
typedef __attribute__((address_space(4))) const char cchar;
__device__ __forceinline__ uint4 sload16(const char *pyr, uint32_t off) {
    const u32x4 v = *(const __attribute__((address_space(4))) u32x4 *)((cchar *)pyr + off);
    return make_uint4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ void axis_terms(float p, int s, uint32_t mul, uint32_t sub, uint32_t t[2], float w[2]) {
    const float u = p * (float)s - 0.5f;
    const float fl = __builtin_floorf(u);
    const int i0 = (int)fl;
    const float f = u - fl;
    const int i1 = i0 + 1;
    w[0] = (i0 >= 0) ? 1.0f - f : 0.0f;  // i0 <= s - 1 always (p <= 1)
    w[1] = (i1 < s) ? f : 0.0f;          // i1 >= 0 always (p >= 0)
    const int c0 = i0 < 0 ? 0 : i0, c1 = i1 > s - 1 ? s - 1 : i1;
    // brick index < 2^7 and stride < 2^24 (dim <= 512): full-rate 24-bit multiplies
    t[0] = __umul24((uint32_t)(c0 >> 2), mul) + (uint32_t)(c0 & 3) * sub;
    t[1] = __umul24((uint32_t)(c1 >> 2), mul) + (uint32_t)(c1 & 3) * sub;
}


// All 2x2x2 texels of a uniform fetch are +0 (empty space): their weighted sum
// adds exactly 0, so the weighting is skipped.  The test runs on the scalar
// unit, beside the vector work of the other wavefronts.
__device__ __forceinline__ bool zero16(const uint4 v[4]) {
    uint32_t o = 0u;
#pragma unroll
    for (int c = 0; c < 4; ++c) o |= (v[c].x | v[c].y) | (v[c].z | v[c].w);
    return o == 0u;
}

// The fp16-pair sampler's axis terms: the border bricks hold texels -1 and s
// as zeros, so neither index is clamped nor weight masked.  Indices are
// shifted by one brick (+4, never negative for the 24-bit multiplies) and
// offsets are taken from the level's storage start.
__device__ __forceinline__ void axis_terms_pad(float p, int s, uint32_t mul, uint32_t sub, uint32_t t[2], float w[2]) {
    const float u = p * (float)s - 0.5f;
    const float fl = __builtin_floorf(u);
    const int i0 = (int)fl + 4;  // texel -1 .. s-1, shifted
    const float f = u - fl;
    w[0] = 1.0f - f;
    w[1] = f;
    t[0] = __umul24((uint32_t)(i0 >> 2), mul) + (uint32_t)(i0 & 3) * sub;
    // texel i0 + 1: the next one in the brick, or the first of the next brick
    t[1] = t[0] + ((i0 & 3) == 3 ? mul - 3u * sub : sub);
}

__device__ __forceinline__ float hlo(uint32_t v) { return __half2float(__ushort_as_half((uint16_t)(v & 0xffffu))); }
__device__ __forceinline__ float hhi(uint32_t v) { return __half2float(__ushort_as_half((uint16_t)(v >> 16))); }

// acc += wa * (first texel of the pair) + wb * (second texel): FP32 weights,
// one v_fma_mix_f32 per channel and texel
__device__ __forceinline__ void acc_pair(float4 &acc, uint4 v, float wa, float wb) {
    acc.x = fmaf(hlo(v.x), wa, acc.x);
    acc.y = fmaf(hhi(v.x), wa, acc.y);
    acc.z = fmaf(hlo(v.y), wa, acc.z);
    acc.w = fmaf(hhi(v.y), wa, acc.w);
    acc.x = fmaf(hlo(v.z), wb, acc.x);
    acc.y = fmaf(hhi(v.z), wb, acc.y);
    acc.z = fmaf(hlo(v.w), wb, acc.z);
    acc.w = fmaf(hhi(v.w), wb, acc.w);
}
__device__ __forceinline__ void acc_f4(float4 &acc, uint4 v, float w) {
    acc.x = fmaf(__uint_as_float(v.x), w, acc.x);
    acc.y = fmaf(__uint_as_float(v.y), w, acc.y);
    acc.z = fmaf(__uint_as_float(v.z), w, acc.z);
    acc.w = fmaf(__uint_as_float(v.w), w, acc.w);
}

// acc += lw * (trilinear sample of level L at p): isotropic, or the
// dir^2-weighted mix of the three faces that face the cone (fe: their element
// offsets, 64 * face).  L is wave-uniform.  fmask (wave-uniform) has bit a
// set unless d2[a] is 0 in every lane: a face no lane weights adds exactly 0
// and is not fetched.  Returns (wave-uniform) whether every texel fetched, in
// every lane, was +0 -- reported by the fp16 paths only (ZQ: also test the
// per-lane fetches, which costs vector instructions).
template <int MODE, bool ZQ>
__device__ __forceinline__ bool level_sample(float4 &acc, const char *pyr, __amdgpu_buffer_rsrc_t rs,
                                             const float *__restrict__ scales, const LevelInfo &L, const float p[3],
                                             float lw, const uint32_t fe[3], const float d2[3], uint32_t fmask,
                                             bool fsame) {
    const int s = L.size;
    const bool aniso = L.nfaces == 6;
    uint32_t ty[2], tz[2];
    float wy[2], wz[2];
    if (MODE == 1 || (MODE == 2 && L.fmt == 2)) {  // one texel per element: 8 corners
        axis_terms(p[1], s, L.sy, 4u, ty, wy);
        axis_terms(p[2], s, L.sz, 16u, tz, wz);
        wz[0] *= lw;
        wz[1] *= lw;
        uint32_t tx[2];
        float wx[2];
        axis_terms(p[0], s, L.sx, 1u, tx, wx);
        uint32_t e[8];
        float w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
            e[k] = tx[dx] + ty[dy] + tz[dz];
            w[k] = (wx[dx] * wy[dy]) * wz[dz];
        }
        const int nf = aniso ? 3 : 1;
        for (int a = 0; a < nf; ++a) {
            if (aniso && !((fmask >> a) & 1u)) continue;
            const uint32_t fo = aniso ? fe[a] : 0u;
            const float fw = aniso ? d2[a] : 1.0f;
            if (MODE == 1) {
#pragma unroll
                for (int k = 0; k < 8; ++k) acc_f4(acc, bload16(rs, L.base + (e[k] + fo) * 16u), w[k] * fw);
            } else {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint32_t ee = e[k] + fo;
                    const uint32_t c = bload4(rs, L.base + ee * 4u);
                    const float4 t = decode_fp8x4(c, scales[L.scale + (ee >> 6)]);
                    const float ww = w[k] * fw;
                    acc.x = fmaf(t.x, ww, acc.x);
                    acc.y = fmaf(t.y, ww, acc.y);
                    acc.z = fmaf(t.z, ww, acc.z);
                    acc.w = fmaf(t.w, ww, acc.w);
                }
            }
        }
        return false;
    }
    // fp16 pairs: element x holds texels x and x+1, so one load per (y, z)
    // corner pair; element -1 (border brick) holds (0, texel 0)
    axis_terms_pad(p[1], s, L.sy, 4u, ty, wy);
    axis_terms_pad(p[2], s, L.sz, 16u, tz, wz);
    wz[0] *= lw;
    wz[1] *= lw;
    const float u = p[0] * (float)s - 0.5f;
    const float fl = __builtin_floorf(u);
    const int cx = (int)fl + 4;  // pair element of texel -1 .. s-1, shifted one brick
    const float f = u - fl;
    const float wlo = 1.0f - f, whi = f;
    const uint32_t ex = __umul24((uint32_t)(cx >> 2), L.sx) + (uint32_t)(cx & 3);
    const uint32_t org = L.base - L.pad * 16u;  // the level's storage start
    uint32_t off[4];
    float wa[4], wb[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int dy = c & 1, dz = c >> 1;
        off[c] = org + (ex + ty[dy] + tz[dz]) * 16u;
        const float wyz = wy[dy] * wz[dz];
        wa[c] = wlo * wyz;
        wb[c] = whi * wyz;
    }
    // At 1080p a wavefront's 8x8 pixels span about a voxel, so at the coarser
    // levels all its lanes usually need the same 2x2x2 texels (and faces).  The
    // four corner elements are then fetched once for the wavefront with scalar
    // loads and every lane weights them itself: the same operations on the same
    // values as the per-lane path, without its vector memory traffic.  Corners
    // 0 and 3 together fix all four (the element encoding of a texel index is
    // unique, and corner 3 carries the clamped upper indices).  fsame: the
    // faces of every lane's cone are the first lane's (wave-uniform, decided
    // once per cone).
    const bool same = (!aniso || fsame) && __builtin_amdgcn_readfirstlane(off[0]) == off[0] &&
                      __builtin_amdgcn_readfirstlane(off[3]) == off[3];
    if (__ballot(same) == __ballot(true)) {
        uint32_t so[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) so[c] = __builtin_amdgcn_readfirstlane(off[c]);
        if (!aniso) {
            uint4 v[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = sload16(pyr, so[c]);
            if (zero16(v)) return true;
#pragma unroll
            for (int c = 0; c < 4; ++c) acc_pair(acc, v[c], wa[c], wb[c]);
            return false;
        }
        bool z = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (!((fmask >> a) & 1u)) continue;
            const uint32_t fo = __builtin_amdgcn_readfirstlane(fe[a]) * 16u;
            uint4 v[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = sload16(pyr, so[c] + fo);
            if (zero16(v)) continue;
            z = false;
            float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int c = 0; c < 4; ++c) acc_pair(t, v[c], wa[c], wb[c]);
            acc.x = fmaf(t.x, d2[a], acc.x);
            acc.y = fmaf(t.y, d2[a], acc.y);
            acc.z = fmaf(t.z, d2[a], acc.z);
            acc.w = fmaf(t.w, d2[a], acc.w);
        }
        return z;
    }
    if (!aniso) {
        uint4 v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = bload16(rs, off[c]);
        if (ZQ && __ballot(!zero16(v)) == 0) return true;
#pragma unroll
        for (int c = 0; c < 4; ++c) acc_pair(acc, v[c], wa[c], wb[c]);
        return false;
    }
    bool z = ZQ;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!((fmask >> a) & 1u)) continue;
        uint4 v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = bload16(rs, off[c] + fe[a] * 16u);
        if (ZQ && __ballot(!zero16(v)) == 0) continue;
        z = false;
        float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int c = 0; c < 4; ++c) acc_pair(t, v[c], wa[c], wb[c]);
        acc.x = fmaf(t.x, d2[a], acc.x);
        acc.y = fmaf(t.y, d2[a], acc.y);
        acc.z = fmaf(t.z, d2[a], acc.z);
        acc.w = fmaf(t.w, d2[a], acc.w);
    }
    return z;
}

// textureLod with LINEAR_MIPMAP_LINEAR, the lambda clamp to [0, q] and the
// NEAREST magnification filter, for one wave-uniform level class D:
//   D = -1       lambda <= 0: NEAREST on level 0;
//   0 <= D < q   levels D and D+1 with weights (1 - fr, fr);
//   D = q        lambda >= q (or NaN): level q alone.
__device__ __forceinline__ int level_class(float lambda, int q, float *fr) {
    if (lambda <= 0.0f) {
        *fr = 0.0f;
        return -1;
    }
    if (!(lambda < (float)q)) {
        *fr = 0.0f;
        return q;
    }
    const float fl = __builtin_floorf(lambda);
    int d1 = (int)fl;
    d1 = d1 < 0 ? 0 : (d1 > q - 1 ? q - 1 : d1);
    *fr = lambda - fl;
    return d1;
}

template <int MODE>
__device__ __forceinline__ float4 sample_class(const PyrArgs &pa, __amdgpu_buffer_rsrc_t rs, int D, float fr,
                                               const float p[3], const uint32_t fe[3], const float d2[3],
                                               uint32_t fmask, bool fsame) {
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int q = pa.nlevels - 1;
    if (D < 0) {
        const int dim = pa.dim;
        const int x = (int)__builtin_floorf(p[0] * (float)dim), y = (int)__builtin_floorf(p[1] * (float)dim),
                  z = (int)__builtin_floorf(p[2] * (float)dim);
        if ((unsigned)x >= (unsigned)dim || (unsigned)y >= (unsigned)dim || (unsigned)z >= (unsigned)dim) return acc;
        const LevelInfo &L = pa.lv[0];
        const uint4 v = bload16(rs, L.base + elem_index(L, x, y, z, 0) * 16u);
        if (MODE == 1) return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                                          __uint_as_float(v.w));
        return unpack_h4(v.x, v.y);
    }
    if (D >= q) {
        level_sample<MODE, false>(acc, pa.pyr, rs, pa.scales, pa.lv[q], p, 1.0f, fe, d2, fmask, fsame);
        return acc;
    }
    // The coarser level first.  Its 2x2x2 cell covers the finer level's cell
    // (texels j, j+1 filter texels 2j .. 2j+3, and the finer cell around the
    // same point lies in that range), and in a nested pyramid a filtered texel
    // is zero only where every texel it filters is.  So where the coarser
    // fetch was all zero in every lane, the finer sample adds exactly 0 and is
    // skipped: cones marching through empty space.
    const bool z = level_sample<MODE, true>(acc, pa.pyr, rs, pa.scales, pa.lv[D + 1], p, fr, fe, d2, fmask, fsame);
    if (!(z && pa.nested))
        level_sample<MODE, false>(acc, pa.pyr, rs, pa.scales, pa.lv[D], p, 1.0f - fr, fe, d2, fmask, fsame);
    return acc;
}

// traceCone (phong.frag:135-180).  ULOD: every lane of the wavefront traces a
// cone of the same aperture and initial height, so the step schedule h_i (and
// with it the sampled level) depends only on the step index and the level
// class is read from the first active lane -- or, with `sched` (the diffuse
// cones, whose schedule the host computed for the launch), read from the
// kernel arguments.  Otherwise the lanes' level classes are visited one at a
// time (each pass with a uniform class).
// warp: warpTexture (phong.frag:150-152), each sample position goes through the
// warp texture before the pyramid is sampled (the level stays the unwarped one).
template <int MODE, bool ANISO, bool ULOD>
__device__ float4 trace_cone(const PyrArgs &pa, __amdgpu_buffer_rsrc_t rs, const float pos[3], const float nrm[3],
                             const float dir_in[3], int steps, float bias, float tan_half, float h0,
                             float lod_offset, const float4 *warp, const ConeSchedule *sched = nullptr) {
    float dir[3];
    {
        const float il = __builtin_amdgcn_rsqf(dot3(dir_in, dir_in));
        dir[0] = dir_in[0] * il;
        dir[1] = dir_in[1] * il;
        dir[2] = dir_in[2] * il;
    }
    uint32_t fe[3] = {0u, 0u, 0u};
    float d2[3] = {0.f, 0.f, 0.f};
    uint32_t fmask = 0u;
    bool fsame = true;
    if (ANISO) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            fe[k] = 64u * (uint32_t)(2 * k + (dir[k] >= 0.0f ? 0 : 1));
            d2[k] = dir[k] * dir[k];
            // cones along a grid axis (faces of axis-aligned walls) weight only
            // one or two of the three faces
            if (__ballot(d2[k] != 0.0f) != 0) fmask |= 1u << k;
        }
        const uint32_t fk = fe[0] | (fe[1] << 10) | (fe[2] << 20);
        fsame = __ballot(__builtin_amdgcn_readfirstlane(fk) != fk) == 0;
    }
    const int q = pa.nlevels - 1;
    float c0 = 0.f, c1 = 0.f, c2 = 0.f, a = 0.f;
    const float scale = 1.0f / (float)pa.dim;
    float start[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) start[k] = pos[k] + (bias * nrm[k]) * scale;
    if (ULOD && sched != nullptr && sched->n == steps) {
        // dim is a power of two, so (h * dir) * scale == h * (dir * scale) exactly
        float ds[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) ds[k] = dir[k] * scale;
        for (int i = 0; i < steps && a < 0.95f; ++i) {
            const float h = sched->h[i];
            float sp[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) sp[k] = start[k] + h * ds[k];
            if (sp[0] < 0.f || sp[0] > 1.f || sp[1] < 0.f || sp[1] > 1.f || sp[2] < 0.f || sp[2] > 1.f) break;
            if (warp) warp_sample(warp, sp, sp);
            const float4 smp = sample_class<MODE>(pa, rs, sched->cls[i], sched->fr[i], sp, fe, d2, fmask, fsame);
            const float aa = 1.0f - a;
            c0 += smp.x * aa;
            c1 += smp.y * aa;
            c2 += smp.z * aa;
            a += aa * smp.w;
        }
        return make_float4(c0, c1, c2, a);
    }
    float h = h0;
    for (int i = 0; i < steps && a < 0.95f; ++i) {
        const float r = h * tan_half;
        const float lod = __builtin_amdgcn_logf(fmaxf(1.0f, 2.0f * r));  // v_log_f32: argument >= 1, no denormals
        float sp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) sp[k] = start[k] + (h * dir[k]) * scale;
        if (sp[0] < 0.f || sp[0] > 1.f || sp[1] < 0.f || sp[1] > 1.f || sp[2] < 0.f || sp[2] > 1.f) break;
        if (warp) warp_sample(warp, sp, sp);
        float fr;
        const int D = level_class(lod + lod_offset, q, &fr);
        float4 smp;
        if (ULOD) {
            const int Du = __builtin_amdgcn_readfirstlane(D);
            const float fru = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(fr)));
            smp = sample_class<MODE>(pa, rs, Du, fru, sp, fe, d2, fmask, fsame);
        } else {
            smp = make_float4(0.f, 0.f, 0.f, 0.f);
            bool pending = true;
            while (true) {
                const uint64_t m = __ballot(pending);
                if (m == 0) break;
                const int Du = __builtin_amdgcn_readlane(D, __builtin_ctzll(m));
                if (pending && D == Du) {
                    smp = sample_class<MODE>(pa, rs, Du, fr, sp, fe, d2, fmask, fsame);
                    pending = false;
                }
            }
        }
        const float aa = 1.0f - a;
        c0 += smp.x * aa;
        c1 += smp.y * aa;
        c2 += smp.z * aa;
        a += aa * smp.w;
        h += r;
    }
    return make_float4(c0, c1, c2, a);
}

static __constant__ float kConeDirs[6][3] = {{0.0f, 1.0f, 0.0f},
                                      {0.0f, 0.5f, 0.866025f},
                                      {0.823639f, 0.5f, 0.267617f},
                                      {0.509037f, 0.5f, -0.700629f},
                                      {-0.5909037f, 0.5f, -0.700629f},  // sic: phong.frag:475
                                      {-0.823639f, 0.5f, 0.267617f}};
static __constant__ float kConeW[6] = {0.25f, 0.15f, 0.15f, 0.15f, 0.15f, 0.15f};

// One pixel of the phong.frag gather: 6 diffuse cones (+ the specular cone).
template <int MODE, bool ANISO>
__device__ __forceinline__ void trace_pixel(const TraceLaunch &t, const PyrArgs &pa, __amdgpu_buffer_rsrc_t rs, int px,
                                            int py, bool reflections) {
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
        const float4 r = trace_cone<MODE, ANISO, true>(pa, rs, pos, N, d, t.d_steps, t.d_bias, t.d_tan_half, t.d_h0,
                                                       t.d_lod_offset, t.warp, &t.d_sched);
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
        const float th = tanf(angle / 2.0f);
        // One material under the whole wavefront (the usual case): every lane's
        // cone has the same aperture, so the step schedule and the level
        // classes are the same in every lane.
        const float th0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(th)));
        float4 r;
        if (__ballot(__float_as_int(th) != __float_as_int(th0)) == 0)
            r = trace_cone<MODE, ANISO, true>(pa, rs, pos, N, R, t.s_steps, t.s_bias, th, t.s_h0, t.s_lod_offset,
                                              t.warp);
        else
            r = trace_cone<MODE, ANISO, false>(pa, rs, pos, N, R, t.s_steps, t.s_bias, th, t.s_h0, t.s_lod_offset,
                                               t.warp);
        t.out_spec[pi] = r;
    }
}

inline PyrArgs pyr_args(const PyrDesc &pyr, int nested) {
    PyrArgs a;
    a.pyr = pyr.pyr;
    a.bytes = pyr.bytes;
    a.scales = pyr.scales;
    a.lv = pyr.d_lv;
    a.dim = pyr.dim;
    a.nlevels = pyr.nlevels;
    a.nested = nested;
    return a;
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t pyr_rsrc(const PyrArgs &pa) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)pa.pyr, (short)0, (int)pa.bytes, 0x00020000);
}

}  // namespace vctk
