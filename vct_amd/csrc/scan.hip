// scan.hip -- item offsets for the binned rasterizer, chunk starts, fills.
#include <hipcub/hipcub.hpp>

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

// chunk[c] = largest slot t with offs[t] <= 64*c  (the slot holding item 64c)
__global__ void k_chunk_starts(const uint64_t *__restrict__ offs, int nslots, uint32_t *chunk, size_t cap) {
    const uint64_t total = offs[nslots];
    uint64_t nchunks = (total + 63) >> 6;
    if (nchunks > cap) nchunks = cap;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nchunks;
         c += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = c << 6;
        int lo = 0, hi = nslots - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (offs[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        chunk[c] = (uint32_t)lo;
    }
}

size_t raster_scan_tmp_bytes(size_t nslots) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                     (int)(nslots + 1));
    return bytes;
}

static hipError_t scan_counts(RasterScratch &rs, uint64_t *counts, uint64_t *offs, uint32_t *chunk, size_t nslots,
                              hipStream_t s) {
    // counts[nslots] is kept at 0 so offs[nslots] is the total
    hipError_t e = launch_fill_u64(counts + nslots, 0, 1, s);
    if (e != hipSuccess) return e;
    size_t bytes = rs.scan_tmp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(rs.scan_tmp, bytes, counts, offs, (int)(nslots + 1), s);
    if (e != hipSuccess) return e;
    // each chunk is a chain of dependent loads (the search): 8 wavefronts per
    // SIMD of the 256 CUs keep the most of them in flight
    hipLaunchKernelGGL(k_chunk_starts, dim3(2048), dim3(256), 0, s, offs, (int)nslots, chunk, rs.chunk_cap);
    return hipGetLastError();
}

hipError_t raster_scan(RasterScratch &rs, size_t nslots, hipStream_t s) {
    return scan_counts(rs, rs.counts, rs.offs, rs.chunk, nslots, s);
}

hipError_t raster_scan_blocks(RasterScratch &rs, size_t nslots, hipStream_t s) {
    return scan_counts(rs, rs.bcounts, rs.boffs, rs.bchunk, nslots, s);
}

}  // namespace vctk
