// warp.hip -- the warp texture of src/Application.cpp:235-575, built on the
// GPU every frame it is used.  The reference voxelizes the scene into a 32^3
// occupancy grid, reads it back to the host for per-axis prefix sums and a
// weight table, and renders the 32^3 RGBA16 warpmap with generateWarpmap.frag
// (through the rgba16f weights texture of generateWarpmapWeights.frag).  Here
// the occupancy pass is the voxelizer in occupancy mode (voxelize.hip), then
//   k_warp_partials  one thread per row: inclusive prefix sums along x, y, z
//   k_warpmap        one thread per texel: generateWarpmap.frag:49-99 at the
//                    texel centre, stored as unorm16 values k / 65535
// so there is no host round trip.  Operation order and rounding follow the
// oracle's or_warpmap.
#pragma clang fp contract(off)

#include "vct_device.h"

namespace vctk {

namespace {
constexpr int D = kWarpDim;

__device__ __forceinline__ int occ_at(const uint32_t *occ, int x, int y, int z) {
    return occ[(z * D + y) * D + x] > 0u ? 1 : 0;
}
}  // namespace

// thread r of axis a: the row through (u, v) with u, v the two other axes
__global__ __launch_bounds__(256) void k_warp_partials(const uint32_t *__restrict__ occ, int4 *part) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * D * D) return;
    const int a = t / (D * D), r = t % (D * D), u = r % D, v = r / D;
    int s = 0;
    for (int i = 0; i < D; ++i) {
        int x, y, z;
        if (a == 0) {
            x = i;
            y = u;
            z = v;
        } else if (a == 1) {
            x = u;
            y = i;
            z = v;
        } else {
            x = u;
            y = v;
            z = i;
        }
        s += occ_at(occ, x, y, z);
        int *p = (int *)&part[(z * D + y) * D + x];
        p[a] = s;
    }
}

// warpWeights[0 or 1][n] of Application.cpp:346-370 for n occupied cells of D
__device__ __forceinline__ void warp_weights(int n, float high, float low, float *l, float *h) {
    if (n == 0 || n == D) {
        *l = *h = 1.0f;
        return;
    }
    const int empty = D - n;
    float hh = high;
    float ll = ((float)D - hh * (float)n) / (float)empty;
    if (ll < low) {
        ll = low;
        hh = ((float)D - ll * (float)empty) / (float)n;
    }
    *l = ll;
    *h = hh;
}

__device__ __forceinline__ float unorm16(float v) {
    return (float)(int)(fminf(fmaxf(v, 0.0f), 1.0f) * 65535.0f + 0.5f) / 65535.0f;
}

__global__ __launch_bounds__(256) void k_warpmap(const uint32_t *__restrict__ occ, const int4 *__restrict__ part,
                                                 vct_warpmap_params p, float4 *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D * D * D) return;
    const int x = t % D, y = (t / D) % D, z = t / (D * D);
    const float tc[3] = {((float)x + 0.5f) / (float)D, ((float)y + 0.5f) / (float)D, ((float)z + 0.5f) / (float)D};
    const int o = occ_at(occ, x, y, z);
    const int4 pc = part[t];
    const int psum[3] = {pc.x, pc.y, pc.z};
    const int total[3] = {part[(z * D + y) * D + D - 1].x, part[(z * D + D - 1) * D + x].y,
                          part[((D - 1) * D + y) * D + x].z};
    float r[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float lin = tc[k] * (float)D;
        const float cell = __builtin_floorf(lin), inner = lin - cell;
        float lo, hi;
        warp_weights(total[k], p.high_resolution, p.low_resolution, &lo, &hi);
        if (p.use_weights_texture) {  // through the rgba16f weights texture
            lo = q16(lo);
            hi = q16(hi);
        }
        const float res = o ? hi : lo;
        const float prev = (float)(o ? psum[k] - 1 : psum[k]);
        const float off = lo * (cell - prev) + hi * prev;
        const float inoff = inner * res;
        const float w = (off + inoff) / (float)D;
        r[k] = unorm16((!p.linear && p.axes[k]) ? w : tc[k]);
    }
    const uint32_t bits = ((uint32_t)total[0] & 0x1Fu) | (((uint32_t)total[1] & 0x1Fu) << 5) |
                          (((uint32_t)total[2] & 0x1Fu) << 10) | ((uint32_t)o << 15);
    out[t] = make_float4(r[0], r[1], r[2], (float)bits / 65535.0f);
}

hipError_t launch_warpmap(const uint32_t *occ, int4 *part, const vct_warpmap_params &p, float4 *out,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_warp_partials, dim3((3 * D * D + 255) / 256), dim3(256), 0, s, occ, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_warpmap, dim3((D * D * D + 255) / 256), dim3(256), 0, s, occ, part, p, out);
    return hipGetLastError();
}

}  // namespace vctk
