// raster_common.h -- the binned triangle rasterizer shared by the shadow-map,
// voxelize and G-buffer passes.
//
// Rules (the GL 4.5 rasterization rules as the oracle restates them):
// window coordinates snapped to 1/256 px (round to nearest even), integer edge
// functions, top-left fill rule, zero-area triangles dropped, samples at the
// pixel centre or the standard 4x pattern, varyings at the pixel centre with
// lambda_i = E_i(centre) / area.
//
// Work distribution: a setup kernel writes one 64-byte record per triangle
// slot plus the number of bbox pixels it may cover; an exclusive scan turns
// the counts into item offsets; each wavefront then takes 64 consecutive
// (slot, pixel) items, so a triangle of any size is spread over as many
// wavefronts as it needs and tiny triangles are packed 64 to a wavefront.
#pragma once

#include "vct_device.h"

namespace vctk {

__device__ __forceinline__ float clamp_win(float v) {
    return v < -524288.0f ? -524288.0f : (v > 524288.0f ? 524288.0f : v);
}

// Record layout (4 x int4):
//  r0 = (A0, B0, A1, B1)
//  r1 = (A2, B2, x0 | y0 << 16, bw | bias_bits << 16 | small << 19)
//  r2 = (C0.lo, C0.hi, C1.lo, C1.hi)      or, small: (E0_0, E0_1, E0_2, 0)
//  r3 = (C2.lo, C2.hi, float area, user)  or, small: (0, 0, float area, user)
// "small" triangles (bbox <= 63 px per side, |A|, |B| < 2^14) keep each edge
// function relative to the bbox corner (x0, y0): E_i = E0_i + A_i dx + B_i dy
// with dx, dy < 2^14 in 1/256 px, every term below 2^29 -- the same integers as
// the 64-bit form, evaluated with 24-bit multiply-adds.
// Returns the number of bbox pixels (0 = no fragments).
__device__ __forceinline__ uint64_t tri_setup(const float win[3][2], int W, int H, int cull, int user,
                                              int4 *rec) {
    int64_t X[3], Y[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        X[i] = (int64_t)__builtin_rintf(clamp_win(win[i][0]) * 256.0f);
        Y[i] = (int64_t)__builtin_rintf(clamp_win(win[i][1]) * 256.0f);
    }
    int64_t A[3], B[3], C[3];
    A[0] = Y[1] - Y[2];
    B[0] = X[2] - X[1];
    C[0] = X[1] * Y[2] - X[2] * Y[1];
    A[1] = Y[2] - Y[0];
    B[1] = X[0] - X[2];
    C[1] = X[2] * Y[0] - X[0] * Y[2];
    A[2] = Y[0] - Y[1];
    B[2] = X[1] - X[0];
    C[2] = X[0] * Y[1] - X[1] * Y[0];
    int64_t area = C[0] + C[1] + C[2];
    if (area == 0) return 0;
    if (area < 0) {
        if (cull) return 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            A[i] = -A[i];
            B[i] = -B[i];
            C[i] = -C[i];
        }
        area = -area;
    }
    uint32_t bias_bits = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        bool tl = (A[i] > 0) || (A[i] == 0 && B[i] < 0);
        if (!tl) bias_bits |= 1u << i;
    }
    int64_t mnx = X[0], mxx = X[0], mny = Y[0], mxy = Y[0];
#pragma unroll
    for (int i = 1; i < 3; ++i) {
        mnx = X[i] < mnx ? X[i] : mnx;
        mxx = X[i] > mxx ? X[i] : mxx;
        mny = Y[i] < mny ? Y[i] : mny;
        mxy = Y[i] > mxy ? Y[i] : mxy;
    }
    int64_t x0 = mnx >> 8, x1 = mxx >> 8, y0 = mny >> 8, y1 = mxy >> 8;
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > W - 1) x1 = W - 1;
    if (y1 > H - 1) y1 = H - 1;
    if (x0 > x1 || y0 > y1) return 0;
    uint32_t bw = (uint32_t)(x1 - x0 + 1), bh = (uint32_t)(y1 - y0 + 1);
    int64_t E0[3];
    bool small = bw <= 63 && bh <= 63;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        E0[i] = A[i] * (x0 << 8) + B[i] * (y0 << 8) + C[i];
        const int64_t lim = (int64_t)1 << 14;
        small = small && A[i] > -lim && A[i] < lim && B[i] > -lim && B[i] < lim && E0[i] > -((int64_t)1 << 29) &&
                E0[i] < ((int64_t)1 << 29);
    }
    rec[0] = make_int4((int)A[0], (int)B[0], (int)A[1], (int)B[1]);
    rec[1] = make_int4((int)A[2], (int)B[2], (int)((uint32_t)x0 | ((uint32_t)y0 << 16)),
                       (int)(bw | (bias_bits << 16) | ((small ? 1u : 0u) << 19)));
    if (small) {
        rec[2] = make_int4((int)E0[0], (int)E0[1], (int)E0[2], 0);
        rec[3] = make_int4(0, 0, __float_as_int((float)area), user);
    } else {
        rec[2] = make_int4((int)(uint32_t)(uint64_t)C[0], (int)(uint32_t)((uint64_t)C[0] >> 32),
                           (int)(uint32_t)(uint64_t)C[1], (int)(uint32_t)((uint64_t)C[1] >> 32));
        rec[3] = make_int4((int)(uint32_t)(uint64_t)C[2], (int)(uint32_t)((uint64_t)C[2] >> 32),
                           __float_as_int((float)area), user);
    }
    return (uint64_t)bw * bh;
}

struct Edges {
    int32_t A[3], B[3];  // |A|, |B| <= 2^28 for every window-clamped triangle
    int64_t C[3];        // large triangles
    int32_t E0[3];       // small triangles: edge values at the bbox corner
    int32_t bias[3];
    float area;
    int x0, y0, bw, user;
    bool small;
};

__device__ __forceinline__ void load_edges(const int4 *__restrict__ recs, size_t slot, Edges &e) {
    int4 r0 = recs[slot * 4 + 0], r1 = recs[slot * 4 + 1], r2 = recs[slot * 4 + 2], r3 = recs[slot * 4 + 3];
    e.A[0] = r0.x;
    e.B[0] = r0.y;
    e.A[1] = r0.z;
    e.B[1] = r0.w;
    e.A[2] = r1.x;
    e.B[2] = r1.y;
    e.x0 = r1.z & 0xffff;
    e.y0 = (int)((uint32_t)r1.z >> 16);
    e.bw = r1.w & 0xffff;
    const uint32_t bb = ((uint32_t)r1.w >> 16) & 7u;
    e.small = (((uint32_t)r1.w >> 19) & 1u) != 0u;
    e.bias[0] = (bb & 1u) ? -1 : 0;
    e.bias[1] = (bb & 2u) ? -1 : 0;
    e.bias[2] = (bb & 4u) ? -1 : 0;
    e.E0[0] = r2.x;
    e.E0[1] = r2.y;
    e.E0[2] = r2.z;
    e.C[0] = (int64_t)(((uint64_t)(uint32_t)r2.y << 32) | (uint32_t)r2.x);
    e.C[1] = (int64_t)(((uint64_t)(uint32_t)r2.w << 32) | (uint32_t)r2.z);
    e.C[2] = (int64_t)(((uint64_t)(uint32_t)r3.y << 32) | (uint32_t)r3.x);
    e.area = __int_as_float(r3.z);
    e.user = r3.w;
}

// Edge i at sample (px * 256 + ox, py * 256 + oy), ox, oy in [0, 256).
__device__ __forceinline__ int64_t edge_value(const Edges &e, int i, int px, int py, int ox, int oy) {
    if (e.small) {
        const int dx = ((px - e.x0) << 8) + ox, dy = ((py - e.y0) << 8) + oy;
        return (int64_t)(e.E0[i] + __mul24(e.A[i], dx) + __mul24(e.B[i], dy));
    }
    const int64_t sx = (int64_t)px * 256 + ox, sy = (int64_t)py * 256 + oy;
    return (int64_t)e.A[i] * sx + (int64_t)e.B[i] * sy + e.C[i];
}

__device__ __forceinline__ bool edge_inside(const Edges &e, int px, int py, int ox, int oy) {
    if (e.small) {
        const int dx = ((px - e.x0) << 8) + ox, dy = ((py - e.y0) << 8) + oy;
        bool in = true;
#pragma unroll
        for (int i = 0; i < 3; ++i) in = in && (e.E0[i] + __mul24(e.A[i], dx) + __mul24(e.B[i], dy) + e.bias[i] >= 0);
        return in;
    }
    bool in = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) in = in && (edge_value(e, i, px, py, ox, oy) + e.bias[i] >= 0);
    return in;
}

// lambda_i = E_i / area, each the IEEE quotient (quot_f32 with one FP64
// reciprocal of the area for the three)
__device__ __forceinline__ void edge_bary(const Edges &e, int px, int py, int ox, int oy, float l[3]) {
    const double r = 1.0 / (double)e.area;
    if (e.small) {
        const int dx = ((px - e.x0) << 8) + ox, dy = ((py - e.y0) << 8) + oy;
#pragma unroll
        for (int i = 0; i < 3; ++i) l[i] = quot_f32((float)(e.E0[i] + __mul24(e.A[i], dx) + __mul24(e.B[i], dy)), r);
        return;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) l[i] = quot_f32((float)edge_value(e, i, px, py, ox, oy), r);
}

// 4x MSAA sample offsets in 1/256 px: (0.375,0.125) (0.875,0.375) (0.125,0.625) (0.625,0.875)
__device__ __forceinline__ int msaa_x(int k) { return k == 0 ? 96 : (k == 1 ? 224 : (k == 2 ? 32 : 160)); }
__device__ __forceinline__ int msaa_y(int k) { return k == 0 ? 32 : (k == 1 ? 96 : (k == 2 ? 160 : 224)); }

// Walk every (slot, pixel) item.  F provides item(slot, px, py, edges) and
// finish().  chunk[c] is the slot holding item 64*c for c < chunk_cap.
template <class F>
__global__ __launch_bounds__(256) void k_raster_items(const int4 *__restrict__ recs,
                                                      const uint64_t *__restrict__ offs,
                                                      const uint32_t *__restrict__ chunk, size_t chunk_cap,
                                                      int nslots, F f) {
    const uint64_t total = offs[nslots];
    const uint64_t nchunks = (total + 63) >> 6;
    const int lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t c = wave; c < nchunks; c += nwaves) {
        const uint64_t k = (c << 6) + (uint64_t)lane;
        int lo = c < chunk_cap ? (int)chunk[c] : 0;
        int hi = (c + 1 < nchunks && c + 1 < chunk_cap) ? (int)chunk[c + 1] : nslots - 1;
        if (k < total) {
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (offs[mid] <= k) lo = mid;
                else hi = mid - 1;
            }
            const uint64_t local = k - offs[lo];
            Edges e;
            load_edges(recs, (size_t)lo, e);
            const uint32_t lu = (uint32_t)local;
            const int px = e.x0 + (int)(lu % (uint32_t)e.bw);
            const int py = e.y0 + (int)(lu / (uint32_t)e.bw);
            f.item(lo, px, py, e);
        }
    }
    f.finish();
}

// The same walk for passes whose per-fragment work is heavy (voxelize with
// lighting): F provides begin() (called by every thread before the walk, for
// workgroup-wide set-up; it ends with a barrier if it writes LDS),
// cover(slot, px, py, edges, l[3]) -> bool, shade(slot, l) and finish().  Covered fragments are appended to a per-wavefront queue in
// LDS and shaded 64 at a time, so the shading runs with every lane busy
// instead of with the lanes of uncovered bbox pixels idle.  The queue holds up
// to 127 entries (at most 63 left over plus one chunk of 64).
template <class F>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_raster_queued(const int4 *__restrict__ recs,
                                                       const uint64_t *__restrict__ offs,
                                                       const uint32_t *__restrict__ chunk, size_t chunk_cap,
                                                       int nslots, F f) {
    __shared__ int q_slot[4][128];
    __shared__ float q_l[4][3][128];
    f.begin();
    const uint64_t total = offs[nslots];
    const uint64_t nchunks = (total + 63) >> 6;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    int count = 0;  // wave-uniform queue length
    for (uint64_t c = wave; c < nchunks; c += nwaves) {
        const uint64_t k = (c << 6) + (uint64_t)lane;
        int lo = c < chunk_cap ? (int)chunk[c] : 0;
        int hi = (c + 1 < nchunks && c + 1 < chunk_cap) ? (int)chunk[c + 1] : nslots - 1;
        bool cov = false;
        float l[3];
        if (k < total) {
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (offs[mid] <= k) lo = mid;
                else hi = mid - 1;
            }
            const uint64_t local = k - offs[lo];
            Edges e;
            load_edges(recs, (size_t)lo, e);
            const uint32_t lu = (uint32_t)local;
            const int px = e.x0 + (int)(lu % (uint32_t)e.bw);
            const int py = e.y0 + (int)(lu / (uint32_t)e.bw);
            cov = f.cover(lo, px, py, e, l);
        }
        const uint64_t m = __ballot(cov);
        if (cov) {
            const int pos = count + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            q_slot[w][pos] = lo;
            q_l[w][0][pos] = l[0];
            q_l[w][1][pos] = l[1];
            q_l[w][2][pos] = l[2];
        }
        count += __popcll(m);
        __builtin_amdgcn_wave_barrier();
        if (count >= 64) {
            const int s0 = q_slot[w][lane];
            const float ls[3] = {q_l[w][0][lane], q_l[w][1][lane], q_l[w][2][lane]};
            // move the entries 64 .. count-1 to the front
            const int rest = count - 64;
            int ms = 0;
            float ml0 = 0.f, ml1 = 0.f, ml2 = 0.f;
            if (lane < rest) {
                ms = q_slot[w][64 + lane];
                ml0 = q_l[w][0][64 + lane];
                ml1 = q_l[w][1][64 + lane];
                ml2 = q_l[w][2][64 + lane];
            }
            __builtin_amdgcn_wave_barrier();
            if (lane < rest) {
                q_slot[w][lane] = ms;
                q_l[w][0][lane] = ml0;
                q_l[w][1][lane] = ml1;
                q_l[w][2][lane] = ml2;
            }
            __builtin_amdgcn_wave_barrier();
            count = rest;
            f.shade(s0, ls);
        }
    }
    if (lane < count) {
        const float ls[3] = {q_l[w][0][lane], q_l[w][1][lane], q_l[w][2][lane]};
        f.shade(q_slot[w][lane], ls);
    }
    f.finish();
}

// Grid for k_raster_items: enough resident wavefronts to fill 256 CUs.
constexpr int kRasterBlocks = 2048;

// Workgroups for a walk: as many as the device holds at once, at most
// kRasterBlocks.  The walk splits the chunks evenly over its wavefronts, so
// workgroups beyond what is resident would run as a second round after the
// first, with most of the chip idle (the lit voxelizer holds 5 wavefronts per
// SIMD, not 8).
template <class F, bool QUEUED>
inline int raster_grid() {
    static int per_cu = 0;  // a property of the kernel: computed once
    if (per_cu == 0) {
        int n = 0;
        hipError_t e;
        if constexpr (QUEUED) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_raster_queued<F>, 256, 0);
        else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_raster_items<F>, 256, 0);
        per_cu = (e == hipSuccess && n > 0) ? n : 8;
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        return kRasterBlocks;
    const long g = (long)per_cu * cus;
    return g > kRasterBlocks ? kRasterBlocks : (int)g;
}

}  // namespace vctk
