// learned.hip -- the learned diffuse gather: a 129 -> 64 -> 64 -> 64 -> 3 MLP
// (ReLU hidden layers, BF16 operands on MFMA, FP32 accumulation) that stands
// in for the 6-cone diffuse loop of phong.frag:480-486, trained on the GPU on
// labels the pipeline generates itself.
//
//   k_label64      FP64-accumulated 64-cone reference of the gather: one
//                  wavefront per surface point, one lane per cone.
//   k_features     the network input of a point: 27 trilinear samples of mip
//                  level 2 on a 3x3x3 stencil (one level-2 voxel apart), the
//                  gather axis TBN*(0,1,0) and an 18-term sin/cos encoding of
//                  the texture-space position; padded to 160 BF16 values.
//   k_mlp_train    forward + backward of one 64-row minibatch tile per
//                  workgroup on v_mfma_f32_16x16x32_bf16; per-workgroup
//                  partial gradients.
//   k_adam         sums the partials, Adam update of the FP32 master weights,
//                  re-emits the padded BF16 (and transposed) operand copies.
//   k_learned      inference: one wavefront per 8x8 pixel tile; flags tiles
//                  whose predicted-luminance variance exceeds a threshold for
//                  the plain 6-cone tracer (trace.hip, tile-list mode).
//
// MFMA 16x16x32 BF16 operand layout (lane l of a wavefront): A holds row
// l%16, k = 8*(l/16) .. +7; B holds column l%16, the same k; the FP32 result
// holds rows 4*(l/16) .. +3 of column l%16.
#include <hip/hip_runtime.h>

#include "learned.h"
#include "trace_core.h"

namespace vctk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN stays NaN): v_cvt_pk_bf16_f32
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bf2pack(float a, float b) {
    const bf16x2 v = {(__bf16)a, (__bf16)b};
    return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ uint16_t f2bf(float f) { return (uint16_t)(bf2pack(f, 0.0f) & 0xffffu); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------- sampling
struct PyrView {
    const char *pyr;
    const float *scales;
    const LevelInfo *lv;
    int dim, nlevels;
    // level 2 (the stencil level) by value: its fields then come from the
    // kernel arguments, not from memory that the compiler must reload after
    // every store the gather makes
    LevelInfo l2;
};

// textureLod with LINEAR_MIPMAP_LINEAR on an isotropic pyramid (border 0).
__device__ __forceinline__ float4 sample_iso(const PyrView &v, const float p[3], float lam) {
    const int q = v.nlevels - 1;
    if (!(lam < (float)q)) return level_linear(v.pyr, v.scales, v.lv[q], p);
    if (lam < 0.0f) lam = 0.0f;
    int d1 = (int)floorf(lam);
    d1 = d1 > q - 1 ? q - 1 : d1;
    const float fr = lam - (float)d1;
    const float4 a = level_linear(v.pyr, v.scales, v.lv[d1], p), b = level_linear(v.pyr, v.scales, v.lv[d1 + 1], p);
    return make_float4((1.0f - fr) * a.x + fr * b.x, (1.0f - fr) * a.y + fr * b.y, (1.0f - fr) * a.z + fr * b.z,
                       (1.0f - fr) * a.w + fr * b.w);
}

// ---------------------------------------------------------------- labels
// 64 directions: Fibonacci set uniform in cos(theta) around local +y;
// weights y_i / 32 (sum 1).  Same definition as the oracle's or_label.c.
__device__ __forceinline__ void label_dir(int i, double d[3], double *w) {
    const double ga = 3.14159265358979323846 * (3.0 - 2.2360679774997896964);
    const double y = 1.0 - (i + 0.5) / 64.0;
    const double r = sqrt(1.0 - y * y);
    const double phi = ga * i;
    d[0] = r * cos(phi);
    d[1] = y;
    d[2] = r * sin(phi);
    *w = y / 32.0;
}

__global__ __launch_bounds__(64) void k_label64(const vct_gbuffer_texel *__restrict__ pts, int npts, PyrView pv,
                                                Extent ext, float4 *out) {
    const int dim = pv.dim;
    const int lane = threadIdx.x;
    for (int pi = blockIdx.x; pi < npts; pi += gridDim.x) {
        const float4 *g = (const float4 *)&pts[pi];
        const float4 g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
        if (__float_as_uint(g2.w) == 0u) {
            if (lane == 0) out[pi] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        const float gp[3] = {g0.x, g0.y, g0.z};
        double pos[3], N[3];
        double n2 = (double)g1.x * g1.x + (double)g1.y * g1.y + (double)g1.z * g1.z;
        n2 = sqrt(n2);
        N[0] = g1.x / n2;
        N[1] = g1.y / n2;
        N[2] = g1.z / n2;
        for (int k = 0; k < 3; ++k)
            pos[k] = ((double)gp[k] - ext.center[k] - ext.vmin[k]) / ((double)ext.vmax[k] - ext.vmin[k]);
        double dl[3], w;
        label_dir(lane, dl, &w);
        double d[3] = {g2.x * dl[0] + g3.x * dl[1] + g1.x * dl[2], g2.y * dl[0] + g3.y * dl[1] + g1.y * dl[2],
                       g2.z * dl[0] + g3.z * dl[1] + g1.z * dl[2]};
        const double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int k = 0; k < 3; ++k) d[k] /= dn;
        const double tan_half = 0.17632698070846498;  // tan(10 degrees)
        const double scale = 1.0 / dim;
        double start[3];
        for (int k = 0; k < 3; ++k) start[k] = pos[k] + N[k] * scale;
        double c0 = 0, c1 = 0, c2 = 0, a = 0, h = 1.0;
        for (int i = 0; i < 256 && a < 0.999; ++i) {
            const double r = h * tan_half;
            const double lod = log2(fmax(1.0, 2.0 * r)) + 0.5;
            float sp[3];
            bool outside = false;
            double spd[3];
            for (int k = 0; k < 3; ++k) {
                spd[k] = start[k] + h * d[k] * scale;
                if (spd[k] < 0.0 || spd[k] > 1.0) outside = true;
                sp[k] = (float)spd[k];
            }
            if (outside) break;
            // trilinear sampling in FP32; transmittance and radiance accumulate in FP64
            const float4 s = sample_iso(pv, sp, (float)lod);
            double sa = s.w < 0.f ? 0.0 : (s.w > 1.f ? 1.0 : (double)s.w);
            const double ah = 1.0 - sqrt(1.0 - sa);
            const double tr = 1.0 - a;
            c0 += 0.5 * s.x * tr;
            c1 += 0.5 * s.y * tr;
            c2 += 0.5 * s.z * tr;
            a += tr * ah;
            h += 0.5 * r;
        }
        double v[4] = {w * c0, w * c1, w * c2, w * a};
        for (int off = 32; off > 0; off >>= 1)
            for (int q = 0; q < 4; ++q) v[q] += __shfl_xor(v[q], off, 64);
        if (lane == 0) {
            const double occ = 1.0 - (v[3] < 0 ? 0.0 : (v[3] > 1 ? 1.0 : v[3]));
            out[pi] = make_float4((float)(v[0] * occ), (float)(v[1] * occ), (float)(v[2] * occ), (float)occ);
        }
    }
}

// ---------------------------------------------------------------- features
// texel -> 160 BF16 inputs (129 used).  Layout:
//   [0, 108)   27 stencil samples x RGBA, stencil index (iz+1)*9 + (iy+1)*3 + (ix+1)
//   [108, 111) gather axis normalize(B)
//   [111, 129) for c in x,y,z: for f in 1,2,4: sin(pi f p_c), cos(pi f p_c)
//   [129, 160) zero
// The 27 samples sit one level-2 texel apart, so they share one set of
// trilinear weights and together read the 4x4x4 texel block around the
// point: the block is read once (16 rows of 4 texels) and interpolated
// separably, x, then y, then z.


// Texels x0 .. x0+3 of row (y, z) of face 0 of an isotropic level, 0 outside
// (x0 >= -2), as float4 (any element format).
__device__ __forceinline__ void row4_f32(const PyrView &pv, const LevelInfo &L, int x0, int y, int z, float4 t[4]) {
    const int s = L.size;
#pragma unroll
    for (int k = 0; k < 4; ++k) t[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    if ((unsigned)y >= (unsigned)s || (unsigned)z >= (unsigned)s) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int x = x0 + k;
        if ((unsigned)x < (unsigned)s) t[k] = load_texel(pv.pyr, pv.scales, L, x, y, z, 0);
    }
}

// The same row of an fp16-pair level, kept as fp16: t[k] = (rg, ba) halves.
// The pair of the last texel ends in the zero border.
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
struct h4 {
    h2 rg, ba;
};
__device__ __forceinline__ h4 as_h4(uint32_t a, uint32_t b) {
    h4 r;
    r.rg = __builtin_bit_cast(h2, a);
    r.ba = __builtin_bit_cast(h2, b);
    return r;
}
// Texels x0 .. x0+3 of row (y, z), zero outside the level: pair elements x0
// (texels x0, x0+1) and x0+2.  A block with first texel i0 - 1, -2 <= i0 <= s
// on every axis (always so for a point in [0,1]^3), spans texels -3 .. s+2,
// inside the zero border bricks around the level (element -1 holds
// (0, texel 0)), so no index is clamped and no texel masked; element indices
// below 0 wrap, and the byte offsets are formed in 32 bits.  A block further
// out (in = false) is all zero and is not read.  SCALAR: x0, y, z are the same
// in every active lane, and the elements come through scalar loads (the
// pyramid is not written while the gather runs).
struct RowRaw {
    uint4 a, b;
};
template <bool SCALAR>
__device__ __forceinline__ void row_fetch(const PyrView &pv, const LevelInfo &L, int x0, int y, int z, bool in,
                                          RowRaw &r) {
    if (!in) {
        r.a = r.b = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const uint32_t oa = L.base + elem_index(L, x0, y, z, 0) * 16u, ob = L.base + elem_index(L, x0 + 2, y, z, 0) * 16u;
    if (SCALAR) {
        r.a = sload16(pv.pyr, oa);
        r.b = sload16(pv.pyr, ob);
    } else {
        r.a = *(const uint4 *)(pv.pyr + oa);
        r.b = *(const uint4 *)(pv.pyr + ob);
    }
}
__device__ __forceinline__ void row_select(const RowRaw &r, h4 t[4]) {
    t[0] = as_h4(r.a.x, r.a.y);
    t[1] = as_h4(r.a.z, r.a.w);
    t[2] = as_h4(r.b.x, r.b.y);
    t[3] = as_h4(r.b.z, r.b.w);
}

__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float f) {
    const float g = 1.0f - f;
    return make_float4(a.x * g + b.x * f, a.y * g + b.y * f, a.z * g + b.z * f, a.w * g + b.w * f);
}

// The fp16-pair feature pass in pieces, so that inference can hand each
// 32-input chunk to the network as soon as it is complete.  For fp16 levels the
// x and y passes run on packed fp16 (v_pk_fma_f16: two channels per
// instruction, on the loaded halves as they are; 11 significant bits against
// the 8 the BF16 inputs keep) and the z pass in FP32.
struct FeatState {
    float p[3];     // texture-space position
    int i0[3];      // first texel of the 4x4x4 level-2 block is i0 - 1
    bool in;        // -2 <= i0 <= size on every axis: the block reaches the level
    h2 fx, gx, fy, gy;
    float fz, gz;
};
__device__ __forceinline__ void feat_setup(const float4 g0, const PyrView &pv, const Extent &ext, FeatState &F) {
    const float gp[3] = {g0.x, g0.y, g0.z};
    for (int k = 0; k < 3; ++k) F.p[k] = ((gp[k] - ext.center[k]) - ext.vmin[k]) / (ext.vmax[k] - ext.vmin[k]);
    const int s2 = pv.l2.size;
    float f[3];
    for (int k = 0; k < 3; ++k) {
        const float u = F.p[k] * (float)s2 - 0.5f;
        const float fl = floorf(u);
        F.i0[k] = (int)fl;
        f[k] = u - fl;
    }
    F.in = true;
    for (int k = 0; k < 3; ++k) F.in = F.in && F.i0[k] >= -2 && F.i0[k] <= s2;
    const _Float16 fx = (_Float16)f[0], gx = (_Float16)(1.0f - f[0]);
    const _Float16 fy = (_Float16)f[1], gy = (_Float16)(1.0f - f[1]);
    F.fx = h2{fx, fx};
    F.gx = h2{gx, gx};
    F.fy = h2{fy, fy};
    F.gy = h2{gy, gy};
    F.fz = f[2];
    F.gz = 1.0f - f[2];
}
// z slice zs of the block, interpolated in x and y: ycur[oy * 3 + ox].
// UNI: the block's first texel iu (= i0 - 1) is the same in every active lane.
template <bool UNI = false>
__device__ __forceinline__ void feat_slice(const PyrView &pv, const FeatState &F, int zs, h4 ycur[9],
                                           const int *iu = nullptr) {
    const LevelInfo &L2 = pv.l2;
    const int x0 = UNI ? iu[0] : F.i0[0] - 1, y0 = UNI ? iu[1] : F.i0[1] - 1, z = (UNI ? iu[2] : F.i0[2] - 1) + zs;
    // the rows' loads carry no branches, so they issue ahead of their use
    RowRaw raw[4];
    // UNI: every lane has the same block, so the first lane's `in` is everyone's
    const bool in = UNI ? __builtin_amdgcn_readfirstlane((int)F.in) != 0 : F.in;
#pragma unroll
    for (int ys = 0; ys < 4; ++ys) row_fetch<UNI>(pv, L2, x0, y0 + ys, z, in, raw[ys]);
#pragma unroll
    for (int ys = 0; ys < 4; ++ys) {
        h4 t[4];
        row_select(raw[ys], t);
#pragma unroll
        for (int ox = 0; ox < 3; ++ox) {
            h4 xv;
            xv.rg = __builtin_elementwise_fma(t[ox + 1].rg, F.fx, t[ox].rg * F.gx);
            xv.ba = __builtin_elementwise_fma(t[ox + 1].ba, F.fx, t[ox].ba * F.gx);
            if (ys <= 2) {
                ycur[ys * 3 + ox].rg = xv.rg * F.gy;
                ycur[ys * 3 + ox].ba = xv.ba * F.gy;
            }
            if (ys >= 1) {
                h4 &yv = ycur[(ys - 1) * 3 + ox];
                yv.rg = __builtin_elementwise_fma(xv.rg, F.fy, yv.rg);
                yv.ba = __builtin_elementwise_fma(xv.ba, F.fy, yv.ba);
            }
        }
    }
}
// the nine stencil samples between two slices as 18 BF16x2 dwords (RG, BA)
__device__ __forceinline__ void feat_zpass(const FeatState &F, const h4 yprev[9], const h4 ycur[9], uint32_t d[18]) {
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        // fmaf(a, b, +0) is the rounded product a * b (both factors >= 0), written
        // as an FMA so that it is one v_fma_mix_f32 on the fp16 half
        const float vx = fmaf((float)ycur[j].rg.x, F.fz, fmaf((float)yprev[j].rg.x, F.gz, 0.0f));
        const float vy = fmaf((float)ycur[j].rg.y, F.fz, fmaf((float)yprev[j].rg.y, F.gz, 0.0f));
        const float vz = fmaf((float)ycur[j].ba.x, F.fz, fmaf((float)yprev[j].ba.x, F.gz, 0.0f));
        const float vw = fmaf((float)ycur[j].ba.y, F.fz, fmaf((float)yprev[j].ba.y, F.gz, 0.0f));
        d[2 * j] = bf2pack(vx, vy);
        d[2 * j + 1] = bf2pack(vz, vw);
    }
}
// inputs 108 .. 129 (gather axis, position encoding, one zero) as 11 dwords
__device__ __forceinline__ void feat_tail(const float p[3], const float4 g3, uint32_t d[11]) {
    const float b[3] = {g3.x, g3.y, g3.z};
    const float bl = sqrtf(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    float e[22];
    for (int k = 0; k < 3; ++k) e[k] = bl > 0.f ? b[k] / bl : 0.f;
    int o = 3;
    // sin/cos(pi f p) = v_sin/v_cos of f p / 2 revolutions: absolute error ~1e-6, far below the BF16 step
    for (int c = 0; c < 3; ++c)
        for (int fq = 1; fq <= 4; fq <<= 1) {
            const float rev = 0.5f * (float)fq * p[c];
            e[o++] = __builtin_amdgcn_sinf(rev);
            e[o++] = __builtin_amdgcn_cosf(rev);
        }
    e[21] = 0.0f;
#pragma unroll
    for (int k = 0; k < 11; ++k) d[k] = bf2pack(e[2 * k], e[2 * k + 1]);
}

// Writes the 160 inputs of one G-buffer texel (g0 = position, g3 = B) as 80
// packed BF16 pairs to row32 (global or LDS).  FMT: level 2 stored as fp16
// pairs (0) or any other element format (1, all passes in FP32).
template <int FMT>
__device__ __forceinline__ void make_features(const float4 g0, const float4 g3, const PyrView &pv, const Extent &ext,
                                              uint32_t *row32) {
    if (FMT == 0) {
        FeatState F;
        feat_setup(g0, pv, ext, F);
        h4 yprev[9], ycur[9];
        uint32_t d[18];
        for (int zs = 0; zs < 4; ++zs) {
            feat_slice(pv, F, zs, ycur);
            if (zs >= 1) {  // stencils with iz = zs - 2 from slices zs - 1 and zs
                feat_zpass(F, yprev, ycur, d);
#pragma unroll
                for (int k = 0; k < 18; ++k) row32[(zs - 1) * 18 + k] = d[k];
            }
#pragma unroll
            for (int j = 0; j < 9; ++j) yprev[j] = ycur[j];
        }
        uint32_t e[11];
        feat_tail(F.p, g3, e);
#pragma unroll
        for (int k = 0; k < 11; ++k) row32[54 + k] = e[k];
#pragma unroll
        for (int k = 65; k < kMlpIn / 2; ++k) row32[k] = 0u;
        return;
    }
    float p[3];
    const float gp[3] = {g0.x, g0.y, g0.z};
    for (int k = 0; k < 3; ++k) p[k] = ((gp[k] - ext.center[k]) - ext.vmin[k]) / (ext.vmax[k] - ext.vmin[k]);
    const LevelInfo &L2 = pv.l2;
    const int s2 = L2.size;
    int i0[3];
    float f[3];
    for (int k = 0; k < 3; ++k) {
        const float u = p[k] * (float)s2 - 0.5f;
        const float fl = floorf(u);
        i0[k] = (int)fl;
        f[k] = u - fl;
    }
    float4 yprev[9], ycur[9];
    for (int zs = 0; zs < 4; ++zs) {
        for (int ys = 0; ys < 4; ++ys) {
            float4 t[4];
            row4_f32(pv, L2, i0[0] - 1, i0[1] - 1 + ys, i0[2] - 1 + zs, t);
#pragma unroll
            for (int ox = 0; ox < 3; ++ox) {
                const float4 xv = lerp4(t[ox], t[ox + 1], f[0]);
                if (ys <= 2) {
                    const float g = 1.0f - f[1];
                    ycur[ys * 3 + ox] = make_float4(xv.x * g, xv.y * g, xv.z * g, xv.w * g);
                }
                if (ys >= 1) {
                    float4 &yv = ycur[(ys - 1) * 3 + ox];
                    yv.x += xv.x * f[1];
                    yv.y += xv.y * f[1];
                    yv.z += xv.z * f[1];
                    yv.w += xv.w * f[1];
                }
            }
        }
        if (zs >= 1) {
            const int base = (zs - 1) * 9;
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const float4 v = lerp4(yprev[j], ycur[j], f[2]);
                row32[(base + j) * 2 + 0] = bf2pack(v.x, v.y);
                row32[(base + j) * 2 + 1] = bf2pack(v.z, v.w);
            }
        }
#pragma unroll
        for (int j = 0; j < 9; ++j) yprev[j] = ycur[j];
    }
    uint32_t e[11];
    feat_tail(p, g3, e);
#pragma unroll
    for (int k = 0; k < 11; ++k) row32[54 + k] = e[k];
#pragma unroll
    for (int k = 65; k < kMlpIn / 2; ++k) row32[k] = 0u;
}

template <int FMT>
__global__ __launch_bounds__(256) void k_features(const vct_gbuffer_texel *__restrict__ pts, int npts, PyrView pv,
                                                  Extent ext, uint16_t *X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const float4 *g = (const float4 *)&pts[i];
    make_features<FMT>(g[0], g[3], pv, ext, (uint32_t *)(X + (size_t)i * kMlpIn));
}

// ---------------------------------------------------------------- MFMA tiles
// acc (16x16 tile at rows m0, cols n0) += A[m][k] * B[n][k] over k in [0, K),
// A row-major with leading dimension lda, B "n-major" (each column's k-vector
// contiguous) with leading dimension ldb; both BF16; K a multiple of 32.
__device__ __forceinline__ f32x4 mfma_tile(const uint16_t *A, int lda, const uint16_t *B, int ldb, int m0, int n0,
                                           int K, f32x4 acc) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 8;
    for (int k0 = 0; k0 < K; k0 += 32) {
        const bf16x8 a = *(const bf16x8 *)(A + (size_t)(m0 + r) * lda + k0 + kq);
        const bf16x8 b = *(const bf16x8 *)(B + (size_t)(n0 + r) * ldb + k0 + kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// ---------------------------------------------------------------- training
struct MlpDev {
    const float *master;  // [kMlpParams] FP32 (W1, b1, W2, b2, W3, b3, W4, b4; W row-major out x in)
    const uint16_t *w1;   // [64][160]  BF16, zero padded
    const uint16_t *w2;   // [64][64]
    const uint16_t *w3;   // [64][64]
    const uint16_t *w4;   // [16][64]   (rows 3..15 zero)
    const uint16_t *w2t;  // [64][64]   transposed: [in][out]
    const uint16_t *w3t;  // [64][64]
    const uint16_t *w4t;  // [64][32]   [in][out], out 3..31 zero
    float out_scale;      // network output * out_scale = predicted rgb * occlusion
};

__device__ __forceinline__ const float *bias_ptr(const float *master, int layer) {
    // offsets of b1..b4 in the flat parameter vector
    const int off[4] = {kMlpW1, kMlpW1 + 64 + kMlpW2, kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3,
                        kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3 + 64 + kMlpW4};
    return master + off[layer];
}

__device__ __forceinline__ uint32_t mix_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Shared memory of one training workgroup (64 rows, 4 wavefronts).
struct TrainSmem {
    uint16_t x[64 * kMlpIn];      // inputs, row-major
    uint16_t xt[kMlpIn * 64];     // inputs transposed [k][m]
    uint16_t a[3][64 * 64];       // hidden activations (post-ReLU), row-major
    uint16_t at[3][64 * 64];      // transposed
    float y[64 * 16];             // outputs (fp32, 3 used)
    float dz[64 * 64];            // current layer's gradient, fp32 [m][n]
    uint16_t dzb[64 * 64];        // BF16 row-major [m][n]
    uint16_t dzt[64 * 64];        // BF16 transposed [n][m]
    float red[4];
};

// Writes a 16x16 FP32 tile into LDS as bias + ReLU -> BF16 row-major and
// transposed activations.
__device__ __forceinline__ void store_act(f32x4 acc, int m0, int n0, const float *bias, bool relu, uint16_t *a,
                                         uint16_t *at, int ld) {
    const int lane = threadIdx.x & 63;
    const int n = n0 + (lane & 15), mb = m0 + (lane >> 4) * 4;
    for (int i = 0; i < 4; ++i) {
        float v = acc[i] + bias[n];
        if (relu) v = v > 0.f ? v : 0.f;
        const uint16_t h = f2bf(v);
        a[(mb + i) * ld + n] = h;
        at[n * 64 + mb + i] = h;
    }
}

__global__ __launch_bounds__(256) void k_mlp_train(MlpDev W, const uint16_t *__restrict__ X,
                                                   const float4 *__restrict__ labels, int npts, int batch,
                                                   uint32_t seed, int step, float *partials) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    TrainSmem &S = *reinterpret_cast<TrainSmem *>(smem_raw);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int row0 = blockIdx.x * 64;
    // gather the minibatch rows: row j of the global batch uses sample hash(seed, step, j) % npts
    for (int i = tid; i < 64 * (kMlpIn / 8); i += 256) {
        const int m = i / (kMlpIn / 8), c8 = i % (kMlpIn / 8);
        const uint32_t pick = mix_hash(seed ^ mix_hash((uint32_t)step * 0x9e3779b9u + (uint32_t)(row0 + m))) %
                              (uint32_t)npts;
        const uint4 v = *(const uint4 *)(X + (size_t)pick * kMlpIn + c8 * 8);
        *(uint4 *)(S.x + m * kMlpIn + c8 * 8) = v;
        const uint16_t *h = (const uint16_t *)&v;
        for (int j = 0; j < 8; ++j) S.xt[(c8 * 8 + j) * 64 + m] = h[j];
    }
    __syncthreads();
    // forward: each wavefront computes a 16-row stripe of every layer
    const int m0 = wave * 16;
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_tile(S.x, kMlpIn, W.w1, kMlpIn, m0, nt * 16, kMlpIn, acc);
        store_act(acc, m0, nt * 16, bias_ptr(W.master, 0), true, S.a[0], S.at[0], 64);
    }
    __syncthreads();
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_tile(S.a[0], 64, W.w2, 64, m0, nt * 16, 64, acc);
        store_act(acc, m0, nt * 16, bias_ptr(W.master, 1), true, S.a[1], S.at[1], 64);
    }
    __syncthreads();
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_tile(S.a[1], 64, W.w3, 64, m0, nt * 16, 64, acc);
        store_act(acc, m0, nt * 16, bias_ptr(W.master, 2), true, S.a[2], S.at[2], 64);
    }
    __syncthreads();
    {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_tile(S.a[2], 64, W.w4, 64, m0, 0, 64, acc);
        const int n = lane & 15, mb = m0 + (lane >> 4) * 4;
        const float *b4 = bias_ptr(W.master, 3);
        for (int i = 0; i < 4; ++i) S.y[(mb + i) * 16 + n] = acc[i] + (n < 3 ? b4[n] : 0.f);
    }
    __syncthreads();
    // output gradient of the mean squared error over batch x 3 outputs
    const float inv_n = 2.0f / ((float)batch * 3.0f);
    float lsum = 0.f;
    for (int i = tid; i < 64 * 64; i += 256) {
        const int m = i >> 6, n = i & 63;
        float d = 0.f;
        if (n < 3) {
            const uint32_t pick = mix_hash(seed ^ mix_hash((uint32_t)step * 0x9e3779b9u + (uint32_t)(row0 + m))) %
                                  (uint32_t)npts;
            const float4 t = labels[pick];
            const float tv = (n == 0 ? t.x : (n == 1 ? t.y : t.z)) / W.out_scale;
            const float e = S.y[m * 16 + n] - tv;
            lsum += e * e;
            d = inv_n * e;
        }
        S.dz[i] = d;
        S.dzb[i] = f2bf(d);
        S.dzt[n * 64 + m] = S.dzb[i];
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
    if (lane == 0) S.red[wave] = lsum;
    __syncthreads();
    float *P = partials + (size_t)blockIdx.x * kMlpPartialStride;
    if (tid == 0) P[kMlpParams] = (S.red[0] + S.red[1] + S.red[2] + S.red[3]) / ((float)batch * 3.0f);
    // backward, layers 4 .. 1.  dW[n][k] = sum_m dz[m][n] a[m][k]: MFMA with
    // rows n (from dz transposed), columns k (activations transposed), K = m.
    const int woff[4] = {0, kMlpW1 + 64, kMlpW1 + 64 + kMlpW2 + 64, kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3 + 64};
    const int fan_in[4] = {129, 64, 64, 64};
    const int fan_out[4] = {64, 64, 64, 3};
    for (int layer = 3; layer >= 0; --layer) {
        const uint16_t *actT = layer == 0 ? S.xt : S.at[layer - 1];
        const int K_in = layer == 0 ? kMlpIn : 64;
        const int nrow_tiles = layer == 3 ? 1 : 4;  // output units / 16
        const int ncol_tiles = K_in / 16;
        for (int t = wave; t < nrow_tiles * ncol_tiles; t += 4) {
            const int rt = t / ncol_tiles, ct = t % ncol_tiles;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma_tile(S.dzt, 64, actT, 64, rt * 16, ct * 16, 64, acc);
            const int k = ct * 16 + (lane & 15), nb = rt * 16 + (lane >> 4) * 4;
            for (int i = 0; i < 4; ++i) {
                const int n = nb + i;
                if (n < fan_out[layer] && k < fan_in[layer]) P[woff[layer] + n * fan_in[layer] + k] = acc[i];
            }
        }
        // bias gradient: sum over the 64 rows of the fp32 dz
        for (int n = tid; n < fan_out[layer]; n += 256) {
            float s = 0.f;
            for (int m = 0; m < 64; ++m) s += S.dz[m * 64 + n];
            P[woff[layer] + fan_out[layer] * fan_in[layer] + n] = s;
        }
        if (layer == 0) break;
        __syncthreads();
        // da[m][k] = sum_n dz[m][n] W[n][k]: rows m (dz row-major), columns k
        // (W transposed, [k][n]); then dz_prev = da * (a > 0)
        const uint16_t *wt = layer == 3 ? W.w4t : (layer == 2 ? W.w3t : W.w2t);
        const int Kn = layer == 3 ? 32 : 64;
        f32x4 da[4];
        for (int ct = 0; ct < 4; ++ct) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            da[ct] = mfma_tile(S.dzb, 64, wt, Kn, m0, ct * 16, Kn, acc);
        }
        __syncthreads();  // everyone has read dzb before it is overwritten
        const uint16_t *aprev = S.a[layer - 1];
        for (int ct = 0; ct < 4; ++ct) {
            const int n = ct * 16 + (lane & 15), mb = m0 + (lane >> 4) * 4;
            for (int i = 0; i < 4; ++i) {
                const int m = mb + i;
                const float g = bf2f(aprev[m * 64 + n]) > 0.f ? da[ct][i] : 0.f;
                S.dz[m * 64 + n] = g;
                const uint16_t h = f2bf(g);
                S.dzb[m * 64 + n] = h;
                S.dzt[n * 64 + m] = h;
            }
        }
        __syncthreads();
    }
}

// Sums the per-workgroup partial gradients and applies Adam to the FP32
// master weights; re-emits the BF16 operand copies.
__global__ __launch_bounds__(256) void k_adam(float *master, float *m1, float *m2, const float *partials,
                                              int nparts, float lr, float b1, float b2, float eps, int step,
                                              uint16_t *w1, uint16_t *w2, uint16_t *w3, uint16_t *w4, uint16_t *w2t,
                                              uint16_t *w3t, uint16_t *w4t, float *loss_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == kMlpParams && loss_out) {
        float l = 0.f;
        for (int i = 0; i < nparts; ++i) l += partials[(size_t)i * kMlpPartialStride + kMlpParams];
        loss_out[0] = l;
    }
    if (p >= kMlpParams) return;
    float g = 0.f;
    for (int i = 0; i < nparts; ++i) g += partials[(size_t)i * kMlpPartialStride + p];
    const float m = b1 * m1[p] + (1.f - b1) * g;
    const float v = b2 * m2[p] + (1.f - b2) * g * g;
    m1[p] = m;
    m2[p] = v;
    const float mh = m / (1.f - powf(b1, (float)step));
    const float vh = v / (1.f - powf(b2, (float)step));
    const float w = master[p] - lr * mh / (sqrtf(vh) + eps);
    master[p] = w;
    const uint16_t h = f2bf(w);
    // scatter into the operand copies
    if (p < kMlpW1) {
        const int n = p / 129, k = p % 129;
        w1[n * kMlpIn + k] = h;
    } else if (p >= kMlpW1 + 64 && p < kMlpW1 + 64 + kMlpW2) {
        const int q = p - (kMlpW1 + 64), n = q / 64, k = q % 64;
        w2[n * 64 + k] = h;
        w2t[k * 64 + n] = h;
    } else if (p >= kMlpW1 + 64 + kMlpW2 + 64 && p < kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3) {
        const int q = p - (kMlpW1 + 64 + kMlpW2 + 64), n = q / 64, k = q % 64;
        w3[n * 64 + k] = h;
        w3t[k * 64 + n] = h;
    } else if (p >= kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3 + 64 && p < kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3 + 64 + kMlpW4) {
        const int q = p - (kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3 + 64), n = q / 64, k = q % 64;
        w4[n * 64 + k] = h;
        w4t[k * 32 + n] = h;
    }
}

// ---------------------------------------------------------------- inference
// One wavefront per 8x8 tile (64 pixels = four 16-pixel MFMA tiles).  Each
// lane builds its pixel's inputs in LDS.  The layers run transposed,
// D^T = W X^T: the weights are the A operand (loaded once per 16 output units
// and K step, from L1/L2, and applied to all four pixel tiles) and the pixels
// the B operand, so a lane's four results are four consecutive units of one
// pixel and leave as one 8-byte LDS store.  Row strides of 168 and 72 BF16 (84
// and 36 dwords) make a tile's 16 row reads hit disjoint LDS banks.
constexpr int kXs = 168;  // input row stride (>= kMlpIn)
constexpr int kAs = 72;   // hidden activation row stride (>= 64)

struct InferSmem {
    uint16_t x[64 * kXs];  // inputs, then hidden activations (stride kAs)
    float y[64 * 4];
};

// acc[nt][mt] += W[16 nt .. +15][K] * X[16 mt .. +15][K]^T over K = 32 * nks
template <int NT>
__device__ __forceinline__ void mlp_layer_t(f32x4 (&acc)[NT][4], const uint16_t *W, int ldw, const uint16_t *X,
                                            int ldx, int nks) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 8;
    for (int ks = 0; ks < nks; ++ks) {
        bf16x8 b[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) b[mt] = *(const bf16x8 *)(X + (mt * 16 + r) * ldx + ks * 32 + kq);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const bf16x8 a = *(const bf16x8 *)(W + (size_t)(nt * 16 + r) * ldw + ks * 32 + kq);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
                acc[nt][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[mt], acc[nt][mt], 0, 0, 0);
        }
    }
}

// The A fragments of k-step ks of W (rows 16 nt + lane % 16), loaded ahead of
// the layer that uses them.
template <int NT>
__device__ __forceinline__ void load_wfrag(bf16x8 (&a)[NT], const uint16_t *W, int ldw, int ks) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) a[nt] = *(const bf16x8 *)(W + (size_t)(nt * 16 + r) * ldw + ks * 32 + kq);
}

// mlp_layer_t with the fragments of k-step 0 already loaded (a0)
template <int NT>
__device__ __forceinline__ void mlp_layer_pre(f32x4 (&acc)[NT][4], const bf16x8 (&a0)[NT], const uint16_t *W, int ldw,
                                              const uint16_t *X, int ldx, int nks) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 8;
    for (int ks = 0; ks < nks; ++ks) {
        bf16x8 b[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) b[mt] = *(const bf16x8 *)(X + (mt * 16 + r) * ldx + ks * 32 + kq);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const bf16x8 a = ks == 0 ? a0[nt] : *(const bf16x8 *)(W + (size_t)(nt * 16 + r) * ldw + ks * 32 + kq);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
                acc[nt][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[mt], acc[nt][mt], 0, 0, 0);
        }
    }
}

// acc[nt][mt] := the bias of rows 16 nt + 4 (lane / 16) .. +3 (the rows this
// lane's part of the 16x16 FP32 tile holds), so the layer's MFMAs add the
// products onto it.
__device__ __forceinline__ void init_bias_t(f32x4 (&acc)[4][4], const float *bias) {
    const int q = ((threadIdx.x & 63) >> 4) * 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const int n0 = nt * 16 + q;
        const f32x4 b = f32x4{bias[n0], bias[n0 + 1], bias[n0 + 2], bias[n0 + 3]};
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) acc[nt][mt] = b;
    }
}
// ReLU -> BF16 hidden activations [pixel][unit], stride kAs, from accumulators
// that already hold the bias.  ReLU is applied to the rounded BF16 values as a
// signed 16-bit max with 0 (v_pk_max_i16): a negative BF16 value (or -0) has
// its sign bit set, so it becomes +0, a positive one is kept; the same result
// as rounding max(x, 0).
__device__ __forceinline__ uint32_t relu_bf2(float a, float b) {
    uint32_t r;
    // kept as one instruction on the packed pair (the generic vector max makes
    // the compiler unpack the pair again)
    asm("v_pk_max_i16 %0, %1, 0" : "=v"(r) : "v"(bf2pack(a, b)));
    return r;
}
__device__ __forceinline__ void store_relu_t(const f32x4 (&acc)[4][4], uint16_t *out) {
    const int lane = threadIdx.x & 63;
    const int q = (lane >> 4) * 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const int n0 = nt * 16 + q;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int m = mt * 16 + (lane & 15);
            const f32x4 v = acc[nt][mt];
            *(uint2 *)(out + m * kAs + n0) = make_uint2(relu_bf2(v[0], v[1]), relu_bf2(v[2], v[3]));
        }
    }
}

// bias + ReLU -> BF16 hidden activations [pixel][unit], stride kAs
__device__ __forceinline__ void store_hidden_t(const f32x4 (&acc)[4][4], const float *bias, uint16_t *out) {
    const int lane = threadIdx.x & 63;
    const int q = (lane >> 4) * 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const int n0 = nt * 16 + q;
        const float b0 = bias[n0], b1 = bias[n0 + 1], b2 = bias[n0 + 2], b3 = bias[n0 + 3];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int m = mt * 16 + (lane & 15);
            const f32x4 v = acc[nt][mt];
            const uint2 h = make_uint2(bf2pack(fmaxf(v[0] + b0, 0.f), fmaxf(v[1] + b1, 0.f)),
                                       bf2pack(fmaxf(v[2] + b2, 0.f), fmaxf(v[3] + b3, 0.f)));
            *(uint2 *)(out + m * kAs + n0) = h;
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(64) void k_learned(MlpDev W, TraceLaunch t, PyrArgs pa, PyrView pv, float threshold,
                                                uint32_t *fallback_list) {
    constexpr int FMT = MODE == 1 ? 1 : 0;  // level 2 is fp32 in fp32 storage, fp16 pairs otherwise
    __shared__ __align__(16) InferSmem S;
    const vct_gbuffer_texel *__restrict__ gbuf = t.gbuf;
    const int width = t.width, height = t.height;
    const Extent &ext = t.ext;
    const int lane = threadIdx.x;
    const int tiles_x = (width + 7) / 8;
    const int tile = blockIdx.x;
    const int px = (tile % tiles_x) * 8 + (lane & 7), py = (tile / tiles_x) * 8 + (lane >> 3);
    const bool inside = px < width && py < height;
    bool covered = false;
    uint32_t *row32 = (uint32_t *)(S.x + lane * kXs);
    if (inside) {
        const float4 *g = (const float4 *)&gbuf[(size_t)py * width + px];
        const float4 g0 = g[0], g2 = g[2], g3 = g[3];
        covered = __float_as_uint(g2.w) != 0u;
        if (covered) make_features<FMT>(g0, g3, pv, ext, row32);
    }
    if (!covered)
        for (int k = 0; k < kMlpIn / 2; ++k) row32[k] = 0u;
    __syncthreads();
    f32x4 acc[4][4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) acc[nt][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
    mlp_layer_t<4>(acc, W.w1, kMlpIn, S.x, kXs, kMlpIn / 32);
    __syncthreads();  // every lane's input reads are done before the activations overwrite them
    store_hidden_t(acc, bias_ptr(W.master, 0), S.x);
    __syncthreads();
    for (int layer = 1; layer <= 2; ++layer) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) acc[nt][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
        mlp_layer_t<4>(acc, layer == 1 ? W.w2 : W.w3, 64, S.x, kAs, 2);
        __syncthreads();
        store_hidden_t(acc, bias_ptr(W.master, layer), S.x);
        __syncthreads();
    }
    f32x4 o[1][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) o[0][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
    mlp_layer_t<1>(o, W.w4, 64, S.x, kAs, 2);
    if (lane < 16) {  // units 0..3 of pixels 16 mt + lane; units 0..2 are the outputs
        const float *b4 = bias_ptr(W.master, 3);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int m = mt * 16 + lane;
#pragma unroll
            for (int n = 0; n < 3; ++n) S.y[m * 4 + n] = (o[0][mt][n] + b4[n]) * W.out_scale;
        }
    }
    __syncthreads();
    const float r = covered ? S.y[lane * 4 + 0] : 0.f, gch = covered ? S.y[lane * 4 + 1] : 0.f,
                bch = covered ? S.y[lane * 4 + 2] : 0.f;
    // predicted-luminance variance of the tile's covered pixels
    const float lum = 0.2126f * r + 0.7152f * gch + 0.0722f * bch;
    float s1 = covered ? lum : 0.f, s2 = covered ? lum * lum : 0.f, cnt = covered ? 1.f : 0.f;
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_xor(s1, off, 64);
        s2 += __shfl_xor(s2, off, 64);
        cnt += __shfl_xor(cnt, off, 64);
    }
    const float mean = cnt > 0.f ? s1 / cnt : 0.f;
    const float var = cnt > 0.f ? s2 / cnt - mean * mean : 0.f;
    // the butterfly sums are identical on every lane, so the choice is wave-uniform
    if (var > threshold) {  // the tile goes on the list that k_learned_fallback traces
        if (lane == 0) fallback_list[1 + atomicAdd(fallback_list, 1u)] = (uint32_t)tile;
    } else if (inside) {
        t.out_ind[(size_t)py * width + px] = make_float4(r, gch, bch, covered ? 1.0f : 0.0f);
    }
}

// Streamed inference for fp16-pair level 2 (fp16 and fp8 storage).  The 160
// inputs never sit in LDS whole: each 32-input chunk goes to a small buffer as
// soon as the feature pass has produced it and layer 1 takes its MFMA k-step
// on it right away; dwords that run past a chunk wait in registers until the
// buffer is free.  The chunk buffer is then reused for the hidden
// activations.  About 10 KiB of LDS per wavefront instead of 22 KiB: more
// wavefronts per CU to hide the latency of the feature loads and of the
// layer-to-layer LDS round trips.  The inputs, weights and arithmetic are
// those of k_learned.
constexpr int kCs = 40;  // chunk row stride in BF16 (32 inputs + padding)
struct StreamSmem {
    union {
        uint16_t chunk[64 * kCs];
        uint16_t hid[64 * kAs];
    };
    float y[64 * 4];
};

// acc[nt][mt] += W[16 nt .. +15][kw .. kw+31] * X[16 mt .. +15][0 .. 31]^T
__device__ __forceinline__ void mlp_kstep(f32x4 (&acc)[4][4], const uint16_t *W, int ldw, int kw, const uint16_t *X,
                                          int ldx) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 8;
    bf16x8 b[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) b[mt] = *(const bf16x8 *)(X + (mt * 16 + r) * ldx + kq);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const bf16x8 a = *(const bf16x8 *)(W + (size_t)(nt * 16 + r) * ldw + kw + kq);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
            acc[nt][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[mt], acc[nt][mt], 0, 0, 0);
    }
}

// dwords src[0 .. n) of this lane's row to chunk positions [o, o + n)
template <int N>
__device__ __forceinline__ void chunk_put(uint32_t *row, int o, const uint32_t *src) {
#pragma unroll
    for (int k = 0; k < N; ++k) row[o + k] = src[k];
}

template <int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_learned_stream(MlpDev W, TraceLaunch t, PyrArgs pa, PyrView pv,
                                                       float threshold, uint32_t *fallback_list) {
    __shared__ __align__(16) StreamSmem S;
    const vct_gbuffer_texel *__restrict__ gbuf = t.gbuf;
    const int width = t.width, height = t.height;
    const int lane = threadIdx.x;
    const int tiles_x = (width + 7) / 8;
    const int tile = blockIdx.x;
    const int px = (tile % tiles_x) * 8 + (lane & 7), py = (tile / tiles_x) * 8 + (lane >> 3);
    const bool inside = px < width && py < height;
    bool covered = false;
    float4 g0 = make_float4(0.f, 0.f, 0.f, 0.f), g3 = g0;
    if (inside) {
        const float4 *g = (const float4 *)&gbuf[(size_t)py * width + px];
        g0 = g[0];
        g3 = g[3];
        covered = __float_as_uint(g[2].w) != 0u;
    }
    uint32_t *row = (uint32_t *)(S.chunk + lane * kCs);
    f32x4 acc[4][4];
    init_bias_t(acc, bias_ptr(W.master, 0));
    FeatState F;
    if (covered) feat_setup(g0, pv, t.ext, F);
    // Usually every covered pixel of the tile reads the same 4x4x4 level-2
    // block (an 8x8 tile covers about a quarter of a level-2 texel at 1080p):
    // then its rows come through scalar loads and only the weights are per lane.
    int iu[3] = {0, 0, 0};
    bool uni = false;
    {
        const uint64_t cm = __ballot(covered);
        if (cm) {
            const int first = (int)__builtin_ctzll(cm);
#pragma unroll
            for (int k = 0; k < 3; ++k) iu[k] = __builtin_amdgcn_readlane(covered ? F.i0[k] : 0, first) - 1;
            const bool same = !covered || (F.i0[0] - 1 == iu[0] && F.i0[1] - 1 == iu[1] && F.i0[2] - 1 == iu[2]);
            uni = __ballot(!same) == 0;
        }
    }
    h4 yprev[9], ycur[9];
    uint32_t d[18];
    // input dwords 0 .. 53 come from the z passes (18 per slice pair), 54 .. 64
    // from the tail, 65 .. 79 are zero; chunk c holds dwords 16 c .. 16 c + 15
    // (unrolled: the slice registers rotate without copies and the chunk
    // bookkeeping below is resolved at compile time)
#pragma unroll
    for (int zs = 0; zs < 4; ++zs) {
        if (covered) {
            if (uni)
                feat_slice<true>(pv, F, zs, ycur, iu);
            else
                feat_slice<false>(pv, F, zs, ycur);
            if (zs >= 1) feat_zpass(F, yprev, ycur, d);
#pragma unroll
            for (int j = 0; j < 9; ++j) yprev[j] = ycur[j];
        } else {
#pragma unroll
            for (int k = 0; k < 18; ++k) d[k] = 0u;
        }
        if (zs == 0) continue;
        // slice pair c = zs - 1 gives dwords 18 c .. 18 c + 17.  Chunk c already
        // holds its first 2 c dwords (carried from the previous pair); d[0 .. 16 - 2c)
        // complete it and d[16 - 2c .. 18) are the first 2 c + 2 dwords of chunk c + 1.
        const int c = zs - 1;
        if (c == 0) chunk_put<16>(row, 0, d);
        if (c == 1) chunk_put<14>(row, 2, d);
        if (c == 2) chunk_put<12>(row, 4, d);
        __syncthreads();
        mlp_kstep(acc, W.w1, kMlpIn, c * 32, S.chunk, kCs);
        __syncthreads();
        if (c == 0) chunk_put<2>(row, 0, d + 16);
        if (c == 1) chunk_put<4>(row, 0, d + 14);
        if (c == 2) chunk_put<6>(row, 0, d + 12);
    }
    uint32_t e[11];
    if (covered) {
        feat_tail(F.p, g3, e);
    } else {
#pragma unroll
        for (int k = 0; k < 11; ++k) e[k] = 0u;
    }
    chunk_put<10>(row, 6, e);  // dwords 54 .. 63 complete chunk 3
    __syncthreads();
    mlp_kstep(acc, W.w1, kMlpIn, 96, S.chunk, kCs);
    __syncthreads();
    row[0] = e[10];  // dword 64, then the zero padding: chunk 4
#pragma unroll
    for (int k = 1; k < 16; ++k) row[k] = 0u;
    __syncthreads();
    // each layer's first weight fragments are loaded before the barriers and
    // activation stores that precede the layer, so they arrive meanwhile
    bf16x8 wn[4], w4n[1];
    load_wfrag<4>(wn, W.w2, 64, 0);
    mlp_kstep(acc, W.w1, kMlpIn, 128, S.chunk, kCs);
    __syncthreads();  // every lane's chunk reads are done before the activations overwrite them
    store_relu_t(acc, S.hid);
    __syncthreads();
#pragma unroll
    for (int layer = 1; layer <= 2; ++layer) {
        init_bias_t(acc, bias_ptr(W.master, layer));
        mlp_layer_pre<4>(acc, wn, layer == 1 ? W.w2 : W.w3, 64, S.hid, kAs, 2);
        if (layer == 1) load_wfrag<4>(wn, W.w3, 64, 0);
        else load_wfrag<1>(w4n, W.w4, 64, 0);
        __syncthreads();
        store_relu_t(acc, S.hid);
        __syncthreads();
    }
    f32x4 o[1][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) o[0][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
    mlp_layer_pre<1>(o, w4n, W.w4, 64, S.hid, kAs, 2);
    if (lane < 16) {
        const float *b4 = bias_ptr(W.master, 3);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int m = mt * 16 + lane;
#pragma unroll
            for (int n = 0; n < 3; ++n) S.y[m * 4 + n] = (o[0][mt][n] + b4[n]) * W.out_scale;
        }
    }
    __syncthreads();
    const float r = covered ? S.y[lane * 4 + 0] : 0.f, gch = covered ? S.y[lane * 4 + 1] : 0.f,
                bch = covered ? S.y[lane * 4 + 2] : 0.f;
    const float lum = 0.2126f * r + 0.7152f * gch + 0.0722f * bch;
    float s1 = covered ? lum : 0.f, s2 = covered ? lum * lum : 0.f, cnt = covered ? 1.f : 0.f;
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_xor(s1, off, 64);
        s2 += __shfl_xor(s2, off, 64);
        cnt += __shfl_xor(cnt, off, 64);
    }
    const float mean = cnt > 0.f ? s1 / cnt : 0.f;
    const float var = cnt > 0.f ? s2 / cnt - mean * mean : 0.f;
    if (var > threshold) {  // the tile goes on the list that k_learned_fallback traces
        if (lane == 0) fallback_list[1 + atomicAdd(fallback_list, 1u)] = (uint32_t)tile;
    } else if (inside) {
        t.out_ind[(size_t)py * width + px] = make_float4(r, gch, bch, covered ? 1.0f : 0.0f);
    }
}

// ---------------------------------------------------------------- launchers
static PyrView pyr_view(const PyrDesc &pyr) {
    PyrView v;
    v.pyr = pyr.pyr;
    v.scales = pyr.scales;
    v.lv = pyr.d_lv;
    v.dim = pyr.dim;
    v.nlevels = pyr.nlevels;
    v.l2 = pyr.lv[pyr.nlevels > 2 ? 2 : 0];
    return v;
}

hipError_t launch_label64(const PyrDesc &pyr, const vct_gbuffer_texel *pts, int npts, Extent ext, float4 *out,
                          hipStream_t s) {
    const int blocks = npts < 65536 ? (npts > 0 ? npts : 1) : 65536;
    hipLaunchKernelGGL(k_label64, dim3(blocks), dim3(64), 0, s, pts, npts, pyr_view(pyr), ext, out);
    return hipGetLastError();
}

hipError_t launch_features(const PyrDesc &pyr, const vct_gbuffer_texel *pts, int npts, Extent ext, uint16_t *X,
                           hipStream_t s) {
    if (pyr.lv[2].fmt == 0)
        hipLaunchKernelGGL(k_features<0>, dim3((npts + 255) / 256), dim3(256), 0, s, pts, npts, pyr_view(pyr), ext, X);
    else
        hipLaunchKernelGGL(k_features<1>, dim3((npts + 255) / 256), dim3(256), 0, s, pts, npts, pyr_view(pyr), ext, X);
    return hipGetLastError();
}

static MlpDev mlp_dev(const MlpBuffers &B) {
    MlpDev d;
    d.master = B.master;
    d.w1 = B.w1;
    d.w2 = B.w2;
    d.w3 = B.w3;
    d.w4 = B.w4;
    d.w2t = B.w2t;
    d.w3t = B.w3t;
    d.w4t = B.w4t;
    d.out_scale = B.out_scale;
    return d;
}

hipError_t launch_mlp_train_step(MlpBuffers &B, const uint16_t *X, const float4 *labels, int npts, int batch,
                                 uint32_t seed, int step, float lr, float b1, float b2, float eps, float *loss_out,
                                 hipStream_t s) {
    const int nblk = batch / 64;
    static bool smem_attr = false;
    if (!smem_attr) {  // the training workgroup keeps its tile, activations and gradients in LDS (> 64 KiB)
        hipError_t e = hipFuncSetAttribute((const void *)k_mlp_train, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)sizeof(TrainSmem));
        if (e != hipSuccess) return e;
        smem_attr = true;
    }
    hipLaunchKernelGGL(k_mlp_train, dim3(nblk), dim3(256), sizeof(TrainSmem), s, mlp_dev(B), X, labels, npts, batch,
                       seed, step, B.partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, s, B.master, B.m1, B.m2, B.partials,
                       nblk, lr, b1, b2, eps, step, B.w1, B.w2, B.w3, B.w4, B.w2t, B.w3t, B.w4t, loss_out);
    return hipGetLastError();
}

size_t mlp_train_smem_bytes() { return sizeof(TrainSmem); }

// The 6-cone tracer on the tiles the confidence check rejected: list[0] is
// their number, list[1 ..] their indices (8x8 tiles, row-major).  Each
// workgroup takes list entries blockIdx.x, blockIdx.x + gridDim.x, ...
// wavefront w of workgroup b takes list entries 4 b + w, 4 (b + gridDim.x) + w, ...
template <int MODE>
__global__ __launch_bounds__(256) void k_learned_fallback(TraceLaunch t, PyrArgs pa, const uint32_t *__restrict__ list) {
    const uint32_t n = list[0];
    const int tiles_x = (t.width + 7) / 8;
    const int lane = threadIdx.x & 63;
    for (uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += gridDim.x * 4) {
        const int tile = (int)list[1 + i];
        const int px = (tile % tiles_x) * 8 + (lane & 7), py = (tile / tiles_x) * 8 + (lane >> 3);
        if (px < t.width && py < t.height) trace_pixel<MODE, false>(t, pa, pyr_rsrc(pa), px, py, false);
    }
}

// Workgroups (of 4 wavefronts) of the fallback launch: 16 wavefronts per CU
// when many tiles fall back, and an empty list costs only the launch.
constexpr int kFallbackBlocks = 1024;

hipError_t launch_learned(const MlpBuffers &B, const PyrDesc &pyr, const TraceLaunch &t, float threshold,
                          uint32_t *fallback_list, hipStream_t s) {
    const int tiles = ((t.width + 7) / 8) * ((t.height + 7) / 8);
    // var > threshold never holds for an infinite (or NaN) threshold: no tile
    // can be listed, so neither the list nor the tracer is needed
    const bool fallback = threshold < __builtin_huge_valf();
    if (fallback) {
        hipError_t e = launch_fill_u32(fallback_list, 0u, 1, s);
        if (e != hipSuccess) return e;
    }
    const int mode = pyr.storage == VCT_STORE_FP32 ? 1 : (pyr.storage == VCT_STORE_FP8 ? 2 : 0);
    const MlpDev W = mlp_dev(B);
    const PyrArgs pa = pyr_args(pyr, t.nested);
    const PyrView pv = pyr_view(pyr);
    const int fb_blocks = (tiles + 3) / 4;
    const dim3 fb(fb_blocks < kFallbackBlocks ? fb_blocks : kFallbackBlocks);
    if (mode == 0) {
        hipLaunchKernelGGL(k_learned_stream<0>, dim3(tiles), dim3(64), 0, s, W, t, pa, pv, threshold, fallback_list);
        if (fallback) hipLaunchKernelGGL(k_learned_fallback<0>, fb, dim3(256), 0, s, t, pa, fallback_list);
    } else if (mode == 1) {
        hipLaunchKernelGGL(k_learned<1>, dim3(tiles), dim3(64), 0, s, W, t, pa, pv, threshold, fallback_list);
        if (fallback) hipLaunchKernelGGL(k_learned_fallback<1>, fb, dim3(256), 0, s, t, pa, fallback_list);
    } else {
        hipLaunchKernelGGL(k_learned_stream<2>, dim3(tiles), dim3(64), 0, s, W, t, pa, pv, threshold, fallback_list);
        if (fallback) hipLaunchKernelGGL(k_learned_fallback<2>, fb, dim3(256), 0, s, t, pa, fallback_list);
    }
    return hipGetLastError();
}

// Operand copies from the FP32 master vector (used after set_weights).
__global__ void k_emit_operands(const float *master, uint16_t *w1, uint16_t *w2, uint16_t *w3, uint16_t *w4,
                                uint16_t *w2t, uint16_t *w3t, uint16_t *w4t) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const uint16_t h = f2bf(master[p]);
    if (p < kMlpW1) {
        w1[(p / 129) * kMlpIn + p % 129] = h;
    } else if (p >= kMlpW1 + 64 && p < kMlpW1 + 64 + kMlpW2) {
        const int q = p - (kMlpW1 + 64);
        w2[q] = h;
        w2t[(q % 64) * 64 + q / 64] = h;
    } else if (p >= kMlpW1 + 64 + kMlpW2 + 64 && p < kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3) {
        const int q = p - (kMlpW1 + 64 + kMlpW2 + 64);
        w3[q] = h;
        w3t[(q % 64) * 64 + q / 64] = h;
    } else if (p >= kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3 + 64 && p < kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3 + 64 + kMlpW4) {
        const int q = p - (kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3 + 64);
        w4[q] = h;
        w4t[(q % 64) * 32 + q / 64] = h;
    }
}

hipError_t launch_emit_operands(MlpBuffers &B, hipStream_t s) {
    hipLaunchKernelGGL(k_emit_operands, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, B.master, B.w1, B.w2, B.w3,
                       B.w4, B.w2t, B.w3t, B.w4t);
    return hipGetLastError();
}

}  // namespace vctk
