// scan.hip -- item offsets for the binned rasterizer, chunk starts, fills.
#include "vct_device.h"

namespace vctk {

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(uint64_t *p, uint64_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ __launch_bounds__(256) void k_reduce_stats(const uint32_t *partials, int n, uint32_t *sum0, uint32_t *max1,
                                                      uint32_t *sum2) {
    uint32_t a = 0, b = 0, c = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        a += partials[i * 4 + 0];
        b = partials[i * 4 + 1] > b ? partials[i * 4 + 1] : b;
        c += partials[i * 4 + 2];
    }
    __shared__ uint32_t red[3][256];
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    red[2][threadIdx.x] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] = red[1][threadIdx.x + s] > red[1][threadIdx.x] ? red[1][threadIdx.x + s] : red[1][threadIdx.x];
            red[2][threadIdx.x] += red[2][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (sum0) *sum0 = red[0][0];
        if (max1) *max1 = red[1][0];
        if (sum2) *sum2 = red[2][0];
    }
}

hipError_t launch_reduce_stats(const uint32_t *partials, int n, uint32_t *sum0, uint32_t *max1, uint32_t *sum2,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_stats, dim3(1), dim3(256), 0, s, partials, n, sum0, max1, sum2);
    return hipGetLastError();
}

static unsigned fill_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 4096 ? 4096 : (b == 0 ? 1 : b));
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}
hipError_t launch_fill_u64(uint64_t *p, uint64_t v, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3(fill_blocks(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------- raster scan
// The item offsets of the binned rasterizer: counts[0..n) (and, for the passes
// with a block walk, bcounts[0..n)) -> exclusive offsets offs[0..n] (offs[n] =
// total) and the chunk starts chunk[c] = the slot holding item 64 c.  Three
// launches for both arrays together: tile sums (2048 slots per workgroup), one
// workgroup scanning the tile sums, then each tile's offsets and chunk starts.
// The slot holding item 64 c is the one slot t with offs[t] <= 64 c <
// offs[t + 1], so every slot writes the chunk starts that fall in its range.
constexpr int kScanThreads = 256, kScanPer = 8, kScanTile = kScanThreads * kScanPer;

size_t raster_scan_tiles(size_t nslots) { return (nslots + kScanTile - 1) / kScanTile; }

__device__ __forceinline__ void scan_load8(const uint32_t *__restrict__ a, size_t n, size_t base, uint32_t v[8]) {
    if (base + 8 <= n) {
        const uint4 x = *(const uint4 *)(a + base), y = *(const uint4 *)(a + base + 4);
        v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
        v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = base + j < n ? a[base + j] : 0u;
    }
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// inclusive prefix sum over the lanes of a wavefront
__device__ __forceinline__ uint64_t wave_incl_u64(uint64_t x, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tile_sums(const uint32_t *__restrict__ a,
                                                                 const uint32_t *__restrict__ b, size_t n,
                                                                 uint64_t *tsum) {
    __shared__ uint64_t red[2][kScanThreads / 64];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPer;
    uint32_t v[8];
    uint64_t sa = 0, sb = 0;
    scan_load8(a, n, base, v);
#pragma unroll
    for (int j = 0; j < 8; ++j) sa += v[j];
    if (b) {
        scan_load8(b, n, base, v);
#pragma unroll
        for (int j = 0; j < 8; ++j) sb += v[j];
    }
    sa = wave_sum_u64(sa);
    sb = wave_sum_u64(sb);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = sa;
        red[1][w] = sb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t ta = 0, tb = 0;
        for (int i = 0; i < kScanThreads / 64; ++i) {
            ta += red[0][i];
            tb += red[1][i];
        }
        tsum[2 * blockIdx.x] = ta;
        tsum[2 * blockIdx.x + 1] = tb;
    }
}

// One workgroup: exclusive prefixes of the tile sums (both arrays) and the
// totals offs[n], boffs[n].
__global__ __launch_bounds__(1024) void k_scan_tiles(const uint64_t *__restrict__ tsum, int ntiles, uint64_t *tpre,
                                                     uint64_t *end_a, uint64_t *end_b) {
    __shared__ uint64_t wt[2][16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (ntiles + 1023) / 1024;
    const int t0 = t * per, t1 = t0 + per < ntiles ? t0 + per : ntiles;
    uint64_t sa = 0, sb = 0;
    for (int i = t0; i < t1; ++i) {
        sa += tsum[2 * i];
        sb += tsum[2 * i + 1];
    }
    const uint64_t ia = wave_incl_u64(sa, lane), ib = wave_incl_u64(sb, lane);
    if (lane == 63) {
        wt[0][w] = ia;
        wt[1][w] = ib;
    }
    __syncthreads();
    uint64_t pa = ia - sa, pb = ib - sb, ta = 0, tb = 0;
    for (int i = 0; i < 16; ++i) {
        if (i < w) {
            pa += wt[0][i];
            pb += wt[1][i];
        }
        ta += wt[0][i];
        tb += wt[1][i];
    }
    for (int i = t0; i < t1; ++i) {
        tpre[2 * i] = pa;
        tpre[2 * i + 1] = pb;
        pa += tsum[2 * i];
        pb += tsum[2 * i + 1];
    }
    if (t == 0) {
        *end_a = ta;
        if (end_b) *end_b = tb;
    }
}

// A slot whose item range holds more than kInlineChunks chunk starts (a large
// triangle) is put on the workgroup's list, and a wavefront writes its chunk
// starts afterwards, 64 at a time; one thread would otherwise hold up its tile.
constexpr int kInlineChunks = 16, kBigCap = 256;

struct BigList {
    uint32_t slot[kBigCap];  // bit 31: the block array
    uint64_t c0[kBigCap], c1[kBigCap];
    int n;
};

// One array's offsets and chunk starts for the 8 slots of this thread, given
// the thread's exclusive prefix p.
__device__ __forceinline__ void scan_write8(const uint32_t v[8], size_t n, size_t base, uint64_t p, uint64_t *offs,
                                            uint32_t *chunk, size_t cap, uint32_t tag, BigList &big) {
    uint64_t o[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        o[j] = p;
        p += v[j];
    }
    if (base + 8 <= n) {
        ulonglong2 *d = (ulonglong2 *)(offs + base);
#pragma unroll
        for (int j = 0; j < 4; ++j) d[j] = make_ulonglong2(o[2 * j], o[2 * j + 1]);
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (base + j < n) offs[base + j] = o[j];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (v[j] == 0u) continue;
        const uint64_t c0 = (o[j] + 63) >> 6, c1 = (o[j] + v[j] - 1) >> 6;
        if (c0 > c1 || c0 >= cap) continue;  // no chunk starts in this slot's range
        if (c1 - c0 >= (uint64_t)kInlineChunks) {
            const int i = atomicAdd(&big.n, 1);
            if (i < kBigCap) {
                big.slot[i] = (uint32_t)(base + j) | tag;
                big.c0[i] = c0;
                big.c1[i] = c1;
                continue;
            }
        }
        for (uint64_t c = c0; c <= c1 && c < cap; ++c) chunk[c] = (uint32_t)(base + j);
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_offsets(const uint32_t *__restrict__ a,
                                                               const uint32_t *__restrict__ b, size_t n,
                                                               const uint64_t *__restrict__ tpre, uint64_t *offs_a,
                                                               uint64_t *offs_b, uint32_t *chunk_a, uint32_t *chunk_b,
                                                               size_t cap) {
    __shared__ uint64_t wt[2][kScanThreads / 64];
    __shared__ BigList big;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) big.n = 0;
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPer;
    uint32_t va[8], vb[8];
    uint64_t sa = 0, sb = 0;
    scan_load8(a, n, base, va);
#pragma unroll
    for (int j = 0; j < 8; ++j) sa += va[j];
    if (b) {
        scan_load8(b, n, base, vb);
#pragma unroll
        for (int j = 0; j < 8; ++j) sb += vb[j];
    }
    const uint64_t ia = wave_incl_u64(sa, lane), ib = wave_incl_u64(sb, lane);
    if (lane == 63) {
        wt[0][w] = ia;
        wt[1][w] = ib;
    }
    __syncthreads();
    uint64_t pa = tpre[2 * blockIdx.x] + ia - sa, pb = tpre[2 * blockIdx.x + 1] + ib - sb;
    for (int i = 0; i < w; ++i) {
        pa += wt[0][i];
        pb += wt[1][i];
    }
    scan_write8(va, n, base, pa, offs_a, chunk_a, cap, 0u, big);
    if (b) scan_write8(vb, n, base, pb, offs_b, chunk_b, cap, 0x80000000u, big);
    __syncthreads();
    const int nb = big.n < kBigCap ? big.n : kBigCap;
    for (int i = w; i < nb; i += kScanThreads / 64) {
        const uint32_t sl = big.slot[i];
        uint32_t *ch = (sl & 0x80000000u) ? chunk_b : chunk_a;
        const uint64_t c1 = big.c1[i] < cap - 1 ? big.c1[i] : cap - 1;
        for (uint64_t c = big.c0[i] + lane; c <= c1; c += 64) ch[c] = sl & 0x7fffffffu;
    }
}

hipError_t raster_scan(RasterScratch &rs, size_t nslots, bool blocks, hipStream_t s) {
    const size_t ntiles = raster_scan_tiles(nslots);
    if (ntiles > rs.scan_tiles_cap) return hipErrorInvalidValue;
    const uint32_t *b = blocks ? rs.bcounts : nullptr;
    if (ntiles) {
        hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, rs.counts, b, nslots,
                           rs.tsum);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, rs.tsum, (int)ntiles, rs.tsum + 2 * rs.scan_tiles_cap,
                       rs.offs + nslots, blocks ? rs.boffs + nslots : nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !ntiles) return e;
    hipLaunchKernelGGL(k_scan_offsets, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, rs.counts, b, nslots,
                       rs.tsum + 2 * rs.scan_tiles_cap, rs.offs, rs.boffs, rs.chunk, rs.bchunk, rs.chunk_cap);
    return hipGetLastError();
}

}  // namespace vctk
