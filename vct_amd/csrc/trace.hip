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
// The sampler and the per-pixel gather live in trace_core.h.
#include "trace_core.h"

namespace vctk {

template <int MODE, bool ANISO>
__global__ __launch_bounds__(256) void k_trace(TraceLaunch t, PyrArgs pa, int tiles_x, int tiles, int per_xcd) {
    // XCD-aware tile order: workgroup b runs on XCD b % 8; give XCD k the
    // contiguous tile band [k * per_xcd, (k + 1) * per_xcd).
    const int b = blockIdx.x;
    const int tile = (b & 7) * per_xcd + (b >> 3);
    if (tile >= tiles) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = (tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = (tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    if (px >= t.width || py >= t.height) return;
    trace_pixel<MODE, ANISO>(t, pa, pyr_rsrc(pa), px, py, t.enable_reflections != 0);
}

hipError_t launch_trace(const PyrDesc &pyr, const TraceLaunch &t, hipStream_t s) {
    const int tiles_x = (t.width + 15) / 16, tiles_y = (t.height + 15) / 16;
    const int tiles = tiles_x * tiles_y;
    const int per_xcd = (tiles + 7) / 8;
    dim3 grid((unsigned)(per_xcd * 8));
    const int mode = pyr.storage == VCT_STORE_FP32 ? 1 : (pyr.storage == VCT_STORE_FP8 ? 2 : 0);
#define VCT_TRACE_LAUNCH(M, A)                                                                                      \
    hipLaunchKernelGGL((k_trace<M, A>), grid, dim3(256), 0, s, t, pyr_args(pyr, t.nested), tiles_x, tiles, per_xcd)
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
