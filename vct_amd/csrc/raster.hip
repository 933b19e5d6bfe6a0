// raster.hip -- shadow-map depth pass and the camera G-buffer (depth pre-pass
// + per-pixel surface attributes) on the binned rasterizer.
//
// Shadow map: simple.vert draws every actor with projection*view of the light
// (src/Application.cpp:212-233); only depth matters for cone tracing.
// G-buffer: the depth pre-pass (src/Application.cpp:936-965) plus what
// phong.vert hands phong.frag (shaders/phong.vert:36-59): world position,
// normal, and the TBN columns.  Visibility is an order-independent 64-bit
// atomicMin of (depth bits, ~triangle id): the nearest fragment wins and, at
// equal depth, the triangle drawn last wins, as glDepthFunc(GL_EQUAL) shading
// after the pre-pass does.
#pragma clang fp contract(off)

#include "raster_common.h"
#include "texture.h"

namespace vctk {

__device__ __forceinline__ void world_vertex(const DevScene &sc, int t, int v, float w[3]) {
    const float *M = sc.model + (size_t)sc.obj[t] * 16;
    const float *p = sc.pos + (size_t)t * 9 + v * 3;
    float p4[4] = {p[0], p[1], p[2], 1.0f}, o[4];
    mat4_vec4(M, p4, o);
    w[0] = o[0];
    w[1] = o[1];
    w[2] = o[2];
}

// ---------------------------------------------------------------- shadow map
// One walk draws up to kShadowBatch lights' maps: slot j * ntri + t is
// triangle t seen from the walk's j-th light, whose index in the light array
// sits at bits 5j .. 5j+4 of `li` (light indices are below kMaxLights = 32).
// Each map receives exactly the fragments a walk of its light alone gives it
// (depth is an order-independent atomicMin), so the batching does not change
// a texel.
static_assert(kMaxLights <= 32, "light indices are packed in 5 bits");
static_assert(kShadowBatch * 5 <= 64, "a walk's light indices fit one 64-bit argument");

// A triangle whose bounding box holds more than this many pixels is drawn by
// the block walk (k_raster_blocks), the others by the pixel walk.
constexpr uint64_t kBlockPixels = 256;

// The 8x8 pixel blocks (aligned to multiples of 8) that cover the bounding box
// of a setup record r with cnt bounding-box pixels.
__device__ __forceinline__ uint64_t bbox_blocks(const int4 r[4], uint64_t cnt) {
    const int x0 = r[1].z & 0xffff, y0 = (int)((uint32_t)r[1].z >> 16), bw = r[1].w & 0xffff;
    const int bh = (int)(cnt / (uint64_t)bw);
    const uint64_t nbx = (uint64_t)(((x0 + bw - 1) >> 3) - (x0 >> 3) + 1);
    const uint64_t nby = (uint64_t)(((y0 + bh - 1) >> 3) - (y0 >> 3) + 1);
    return nbx * nby;
}

__global__ __launch_bounds__(256) void k_shadow_setup(DevScene sc, const DevLight *__restrict__ lights, uint64_t li,
                                                      int4 *rec, float4 *attr, uint32_t *counts, uint32_t *bcounts) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= sc.ntri) return;
    const int l = (int)((li >> (5u * blockIdx.y)) & 31u);
    const float *ls = lights[l].ls;
    const int N = (int)lights[l].shadow_size;
    const size_t slot = (size_t)blockIdx.y * (size_t)sc.ntri + (size_t)t;
    float clip[3][4];
    bool behind = false;
    for (int v = 0; v < 3; ++v) {
        float w[3];
        world_vertex(sc, t, v, w);
        float w4[4] = {w[0], w[1], w[2], 1.0f};
        mat4_vec4(ls, w4, clip[v]);
        if (clip[v][2] < -clip[v][3] || clip[v][3] <= 0.0f) behind = true;
    }
    uint64_t cnt = 0;
    if (!behind) {
        float half = (float)N * 0.5f;
        float win[3][2], zw[3];
        for (int v = 0; v < 3; ++v) {
            float xd = clip[v][0] / clip[v][3], yd = clip[v][1] / clip[v][3], zd = clip[v][2] / clip[v][3];
            win[v][0] = xd * half + half;
            win[v][1] = yd * half + half;
            zw[v] = zd * 0.5f + 0.5f;
        }
        int4 r[4];
        // back faces culled: GL_CULL_FACE stays enabled from the previous
        // frame's passes when the reference draws its shadow map
        // (Application.cpp:737,942,979); the record's user field is the light
        cnt = tri_setup(win, N, N, 1, l, r);
        if (cnt) {
            for (int i = 0; i < 4; ++i) rec[slot * 4 + i] = r[i];
            attr[slot] = make_float4(zw[0], zw[1], zw[2], 0.0f);
        }
        if (cnt > kBlockPixels) {  // 8x8 blocks of the bounding box instead of its pixels
            counts[slot] = 0;
            bcounts[slot] = (uint32_t)bbox_blocks(r, cnt);
            return;
        }
    }
    counts[slot] = (uint32_t)cnt;
    bcounts[slot] = 0;
}

// The triangle slot of every block of the block walk, one thread per block
// (the search of k_raster_items, run for 64 blocks per wavefront so that the
// dependent loads of many searches are in flight together).  Blocks beyond
// `cap` are left to the walk's own search.
__global__ __launch_bounds__(256) void k_block_slots(const uint64_t *__restrict__ offs,
                                                     const uint32_t *__restrict__ chunk, size_t chunk_cap, int nslots,
                                                     uint32_t *slots, size_t cap) {
    const uint64_t total = offs[nslots];
    const uint64_t n = total < cap ? total : cap;
    const uint64_t nchunks = (total + 63) >> 6;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t c = k >> 6;
        int lo = c < chunk_cap ? (int)chunk[c] : 0;
        int hi = (c + 1 < nchunks && c + 1 < chunk_cap) ? (int)chunk[c + 1] : nslots - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (offs[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        slots[k] = (uint32_t)lo;
    }
}

// The large triangles of a walk: one wavefront per 8x8-pixel block of a
// triangle's bounding box, its lanes the block's pixels (row-major).  Each
// wavefront takes a contiguous run of blocks, so consecutive blocks mostly
// belong to one triangle and its record (wave-uniform) is read again only when
// the run enters the next triangle (its slot from k_block_slots, or the same
// search past that table).  Each lane's edge values are the block corner's
// (wave-uniform) plus the lane's offset A_i dx + B_i dy, set once per triangle:
// in 32-bit wrapping arithmetic for the small form (the same value modulo 2^32
// as edge_value's sum), in FP64 for the 64-bit form (every term is an integer
// below 2^53, so each sum is exact).  The barycentrics are quot_f32(E, area)
// with the triangle's reciprocal, so every fragment and its barycentrics are
// those the pixel walk produces.  F provides enter(slot, edges) (the slot's
// wave-uniform data, and the image size W x H) and frag(edges, px, py, l).
template <class F>
__global__ __launch_bounds__(256) void k_raster_blocks(const int4 *__restrict__ recs, const uint64_t *__restrict__ offs,
                                                       const uint32_t *__restrict__ chunk, size_t chunk_cap, int nslots,
                                                       const uint32_t *__restrict__ bslots, size_t bslots_cap, F fr) {
    F f = fr;
    const uint64_t total = offs[nslots];
    const uint64_t nchunks = (total + 63) >> 6;
    const int lane = threadIdx.x & 63;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t wave = __builtin_amdgcn_readfirstlane((uint32_t)(((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const uint64_t per = (total + nwaves - 1) / nwaves;
    const uint64_t k0 = wave * per, k1 = k0 + per < total ? k0 + per : total;
    uint64_t run_lo = 0, run_hi = 0;  // the blocks [run_lo, run_hi) of the current slot
    int slot = 0;
    Edges e;
    uint32_t nbx = 1, bxi = 0, byi = 0;  // the block's column and row in the triangle's bounding box
    double rcp = 0.0;                    // RN64(1 / area)
    uint32_t off32[3] = {0u, 0u, 0u};    // small form: A_i dx + B_i dy of this lane (mod 2^32)
    double offd[3] = {0.0, 0.0, 0.0}, Ad[3] = {0.0, 0.0, 0.0}, Bd[3] = {0.0, 0.0, 0.0}, Cd[3] = {0.0, 0.0, 0.0};
    const int dxl = (lane & 7) * 256, dyl = (lane >> 3) * 256;
    for (uint64_t k = k0; k < k1; ++k, ++bxi) {
        if (bxi == nbx) {
            bxi = 0;
            ++byi;
        }
        if (k >= run_hi) {  // the run enters the next triangle
            int lo;
            if (k < bslots_cap) {
                lo = (int)bslots[k];
            } else {
                const uint64_t c = k >> 6;  // block k lies in chunk c (64 blocks per chunk)
                lo = c < chunk_cap ? (int)chunk[c] : 0;
                int hi = (c + 1 < nchunks && c + 1 < chunk_cap) ? (int)chunk[c + 1] : nslots - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (offs[mid] <= k) lo = mid;
                    else hi = mid - 1;
                }
            }
            slot = __builtin_amdgcn_readfirstlane(lo);
            run_lo = offs[slot];
            run_hi = offs[slot + 1];
            load_edges(recs, (size_t)slot, e);
            f.enter(slot, e);
            nbx = (uint32_t)(((e.x0 + e.bw - 1) >> 3) - (e.x0 >> 3) + 1);
            const uint32_t local = (uint32_t)(k - run_lo);  // the run may enter a triangle mid-way
            bxi = local % nbx;
            byi = local / nbx;
            rcp = 1.0 / (double)e.area;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                if (e.small) {
                    off32[i] = (uint32_t)e.A[i] * (uint32_t)dxl + (uint32_t)e.B[i] * (uint32_t)dyl;
                } else {
                    Ad[i] = (double)e.A[i];
                    Bd[i] = (double)e.B[i];
                    Cd[i] = (double)e.C[i];
                    offd[i] = Ad[i] * (double)dxl + Bd[i] * (double)dyl;
                }
            }
        }
        const int bx = ((e.x0 >> 3) + (int)bxi) * 8, by = ((e.y0 >> 3) + (int)byi) * 8;
        const int px = bx + (lane & 7), py = by + (lane >> 3);
        if (px >= f.W || py >= f.H) continue;  // the block overhangs the image
        float l[3];
        bool in = true;
        if (e.small) {
            // edge_value's dx = ((px - x0) << 8) + 128 = cx + dxl, likewise dy
            const uint32_t cx = (uint32_t)(((bx - e.x0) << 8) + 128), cy = (uint32_t)(((by - e.y0) << 8) + 128);
            int32_t E[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const uint32_t corner = (uint32_t)e.E0[i] + (uint32_t)e.A[i] * cx + (uint32_t)e.B[i] * cy;
                E[i] = (int32_t)(corner + off32[i]);
                in = in && ((int32_t)((uint32_t)E[i] + (uint32_t)e.bias[i]) >= 0);
            }
            if (!in) continue;
#pragma unroll
            for (int i = 0; i < 3; ++i) l[i] = quot_f32((float)E[i], rcp);
        } else {
            const double sx = (double)(bx * 256 + 128), sy = (double)(by * 256 + 128);
            double E[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                E[i] = fma(Ad[i], sx, fma(Bd[i], sy, Cd[i])) + offd[i];
                in = in && (E[i] >= (double)(-e.bias[i]));
            }
            if (!in) continue;
#pragma unroll
            for (int i = 0; i < 3; ++i) l[i] = quot_f32((float)E[i], rcp);
        }
        f.frag(e, px, py, l);
    }
}

// ShadowFrag for the block walk.
struct ShadowBlockFrag {
    const float4 *attr;
    const DevLight *lights;
    float4 a;
    uint32_t *map;
    int W, H;
    __device__ void enter(int slot, const Edges &e) {
        a = attr[slot];
        const DevLight &L = lights[e.user];
        map = (uint32_t *)const_cast<float *>(L.shadowmap);
        W = H = (int)L.shadow_size;
    }
    __device__ void frag(const Edges &, int px, int py, const float l[3]) {
        float z = interp3(l, a.x, a.y, a.z);
        z = z + 0.0f;
        if (!(z >= 0.0f && z <= 1.0f)) return;
        atomicMin(&map[(size_t)py * W + px], __float_as_uint(z));
    }
};

struct ShadowFrag {
    const float4 *attr;
    const DevLight *lights;
    __device__ void item(int slot, int px, int py, const Edges &e) {
        if (!edge_inside(e, px, py, 128, 128)) return;
        float l[3];
        edge_bary(e, px, py, 128, 128, l);
        float4 a = attr[slot];
        float z = interp3(l, a.x, a.y, a.z);
        z = z + 0.0f;
        if (!(z >= 0.0f && z <= 1.0f)) return;
        const DevLight &L = lights[e.user];
        uint32_t *depth = (uint32_t *)const_cast<float *>(L.shadowmap);
        atomicMin(&depth[(size_t)py * L.shadow_size + px], __float_as_uint(z));
    }
    __device__ void finish() {}
};

__global__ void k_shadow_quads_batch(const DevLight *__restrict__ lights, uint64_t li);

hipError_t launch_shadowmaps(const DevScene &sc, const DevLight *hlights, const DevLight *dlights, const int *li,
                             int n, RasterScratch &rs, hipStream_t s) {
    hipError_t e;
    const int batch = shadow_batch(sc.ntri);
    for (int j0 = 0; j0 < n; j0 += batch) {
        const int nb = n - j0 < batch ? n - j0 : batch;
        uint64_t packed = 0;
        for (int j = 0; j < nb; ++j) {
            const DevLight &L = hlights[li[j0 + j]];
            const size_t N = L.shadow_size;  // clear depth 1.0
            if ((e = launch_fill_u32((uint32_t *)const_cast<float *>(L.shadowmap), 0x3f800000u, N * N, s)) !=
                hipSuccess)
                return e;
            packed |= (uint64_t)(uint32_t)li[j0 + j] << (5 * j);
        }
        if (sc.ntri > 0) {
            const size_t nslots = (size_t)nb * (size_t)sc.ntri;
            hipLaunchKernelGGL(k_shadow_setup, dim3((sc.ntri + 255) / 256, nb), dim3(256), 0, s, sc, dlights, packed,
                               rs.rec, rs.attr, rs.counts, rs.bcounts);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            if ((e = raster_scan(rs, nslots, true, s)) != hipSuccess) return e;
            ShadowFrag f{rs.attr, dlights};
            hipLaunchKernelGGL(k_raster_items<ShadowFrag>, dim3(raster_grid<ShadowFrag, false>()), dim3(256), 0, s,
                               rs.rec, rs.offs, rs.chunk, rs.chunk_cap, (int)nslots, f);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            hipLaunchKernelGGL(k_block_slots, dim3(kRasterBlocks), dim3(256), 0, s, rs.boffs, rs.bchunk, rs.chunk_cap,
                               (int)nslots, rs.bslots, rs.bslots_cap);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            ShadowBlockFrag fb{rs.attr, dlights, make_float4(0.f, 0.f, 0.f, 0.f), nullptr, 0, 0};
            hipLaunchKernelGGL(k_raster_blocks<ShadowBlockFrag>, dim3(kRasterBlocks), dim3(256), 0, s, rs.rec,
                               rs.boffs, rs.bchunk, rs.chunk_cap, (int)nslots, rs.bslots, rs.bslots_cap, fb);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        // the footprint layout of every map of the walk, one launch
        int n1max = 0;
        for (int j = 0; j < nb; ++j) {
            const DevLight &L = hlights[li[j0 + j]];
            const int n1 = (int)L.shadow_size + 1;
            if (L.shadowquad && n1 * n1 > n1max) n1max = n1 * n1;
        }
        if (n1max > 0) {
            hipLaunchKernelGGL(k_shadow_quads_batch, dim3((n1max + 255) / 256, nb), dim3(256), 0, s, dlights, packed);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

__device__ __forceinline__ void shadow_quad(const float *__restrict__ sm, int N, float4 *q, int idx) {
    const int n1 = N + 1;
    const int i = idx % n1 - 1, j = idx / n1 - 1;
    auto tex = [&](int x, int y) -> float {
        return (x >= 0 && x < N && y >= 0 && y < N) ? sm[(size_t)y * N + x] : 1.0f;
    };
    q[idx] = make_float4(tex(i, j), tex(i + 1, j), tex(i, j + 1), tex(i + 1, j + 1));
}

__global__ __launch_bounds__(256) void k_shadow_quads(const float *__restrict__ sm, int N, float4 *q) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (N + 1) * (N + 1)) return;
    shadow_quad(sm, N, q, idx);
}

// The footprint layouts of a shadow walk's maps: blockIdx.y is the walk's
// j-th light (its index at bits 5j .. 5j+4 of li); lights without a layout
// buffer are skipped.
__global__ __launch_bounds__(256) void k_shadow_quads_batch(const DevLight *__restrict__ lights, uint64_t li) {
    const DevLight &L = lights[(li >> (5u * blockIdx.y)) & 31u];
    const int N = (int)L.shadow_size;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (!L.shadowquad || idx >= (N + 1) * (N + 1)) return;
    shadow_quad(L.shadowmap, N, const_cast<float4 *>(L.shadowquad), idx);
}

hipError_t launch_shadow_quads(const float *depth, int N, float4 *quads, hipStream_t s) {
    if (!quads) return hipSuccess;
    const int n = (N + 1) * (N + 1);
    hipLaunchKernelGGL(k_shadow_quads, dim3((n + 255) / 256), dim3(256), 0, s, depth, N, quads);
    return hipGetLastError();
}

// ---------------------------------------------------------------- G-buffer
struct CamMats {
    float pv[16];
    int W, H, cull;
};

__global__ __launch_bounds__(256) void k_gbuf_setup(DevScene sc, CamMats cm, int4 *rec, float4 *attr,
                                                    uint32_t *counts, uint32_t *bcounts) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= sc.ntri) return;
    float clip[3][4];
    for (int v = 0; v < 3; ++v) {
        float w[3];
        world_vertex(sc, t, v, w);
        float w4[4] = {w[0], w[1], w[2], 1.0f};
        mat4_vec4(cm.pv, w4, clip[v]);
    }
    // Sutherland-Hodgman against the near plane z >= -w
    float poly[4][4];
    int n = 0;
    float d[3];
    for (int i = 0; i < 3; ++i) d[i] = clip[i][2] + clip[i][3];
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3;
        bool ii = d[i] >= 0.0f, jj = d[j] >= 0.0f;
        if (ii) {
            for (int k = 0; k < 4; ++k) poly[n][k] = clip[i][k];
            n++;
        }
        if (ii != jj) {
            float tt = d[i] / (d[i] - d[j]);
            for (int k = 0; k < 4; ++k) poly[n][k] = clip[i][k] + tt * (clip[j][k] - clip[i][k]);
            n++;
        }
    }
    const float hw = (float)cm.W * 0.5f, hh = (float)cm.H * 0.5f;
    for (int sub = 0; sub < 2; ++sub) {
        const size_t slot = (size_t)t * 2 + sub;
        uint64_t cnt = 0;
        if (sub + 2 < n) {  // fan (0,1,2) needs 3 vertices, (0,2,3) needs 4
            const int vi[3] = {0, sub + 1, sub + 2};
            float win[3][2], zw[3];
            bool bad = false;
            for (int v = 0; v < 3; ++v) {
                const float *c = poly[vi[v]];
                float w = c[3];
                if (!(w > 0.0f)) bad = true;
                float xd = c[0] / w, yd = c[1] / w, zd = c[2] / w;
                win[v][0] = xd * hw + hw;
                win[v][1] = yd * hh + hh;
                zw[v] = zd * 0.5f + 0.5f;
            }
            if (!bad) {
                int4 r[4];
                cnt = tri_setup(win, cm.W, cm.H, cm.cull, t, r);
                if (cnt) {
                    for (int i = 0; i < 4; ++i) rec[slot * 4 + i] = r[i];
                    attr[slot] = make_float4(zw[0], zw[1], zw[2], 0.0f);
                }
                if (cnt > kBlockPixels) {  // 8x8 blocks of the bounding box instead of its pixels
                    counts[slot] = 0;
                    bcounts[slot] = (uint32_t)bbox_blocks(r, cnt);
                    continue;
                }
            }
        }
        counts[slot] = (uint32_t)cnt;
        bcounts[slot] = 0;
    }
}

struct GbufFrag {
    const float4 *attr;
    unsigned long long *vis;
    int W;
    __device__ void item(int slot, int px, int py, const Edges &e) {
        if (!edge_inside(e, px, py, 128, 128)) return;
        float l[3];
        edge_bary(e, px, py, 128, 128, l);
        float4 a = attr[slot];
        float z = interp3(l, a.x, a.y, a.z);
        z = z + 0.0f;
        if (!(z >= 0.0f && z <= 1.0f)) return;
        const uint32_t tri = (uint32_t)e.user;
        unsigned long long key = ((unsigned long long)__float_as_uint(z) << 32) | (unsigned long long)(0xffffffffu - tri);
        atomicMin(&vis[(size_t)py * W + px], key);
    }
    __device__ void finish() {}
};

// GbufFrag for the block walk.
struct GbufBlockFrag {
    const float4 *attr;
    unsigned long long *vis;
    int W, H;
    float4 a;
    __device__ void enter(int slot, const Edges &) { a = attr[slot]; }
    __device__ void frag(const Edges &e, int px, int py, const float l[3]) {
        float z = interp3(l, a.x, a.y, a.z);
        z = z + 0.0f;
        if (!(z >= 0.0f && z <= 1.0f)) return;
        const uint32_t tri = (uint32_t)e.user;
        const unsigned long long key =
            ((unsigned long long)__float_as_uint(z) << 32) | (unsigned long long)(0xffffffffu - tri);
        atomicMin(&vis[(size_t)py * W + px], key);
    }
};

__global__ __launch_bounds__(256) void k_gbuf_resolve(DevScene sc, CamMats cm, const unsigned long long *vis,
                                                      vct_gbuffer_texel *out) {
    const int px = blockIdx.x * 16 + (threadIdx.x & 15), py = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (px >= cm.W || py >= cm.H) return;
    const size_t pi = (size_t)py * cm.W + px;
    float4 *o = (float4 *)&out[pi];
    const unsigned long long key = vis[pi];
    if (key == ~0ull) {
        float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        o[0] = z;
        o[1] = z;
        o[2] = z;
        o[3] = z;
        return;
    }
    const int t = (int)(0xffffffffu - (uint32_t)(key & 0xffffffffull));
    const float xn = (2.0f * ((float)px + 0.5f)) / (float)cm.W - 1.0f;
    const float yn = (2.0f * ((float)py + 0.5f)) / (float)cm.H - 1.0f;
    float Wp[3][3], clip[3][4];
    for (int v = 0; v < 3; ++v) {
        world_vertex(sc, t, v, Wp[v]);
        float w4[4] = {Wp[v][0], Wp[v][1], Wp[v][2], 1.0f};
        mat4_vec4(cm.pv, w4, clip[v]);
    }
    float ax[3], ay[3];
    for (int v = 0; v < 3; ++v) {
        ax[v] = clip[v][0] - xn * clip[v][3];
        ay[v] = clip[v][1] - yn * clip[v][3];
    }
    float b[3];
    b[0] = ax[1] * ay[2] - ay[1] * ax[2];
    b[1] = ax[2] * ay[0] - ay[2] * ax[0];
    b[2] = ax[0] * ay[1] - ay[0] * ax[1];
    const float sum = (b[0] + b[1]) + b[2];
    const double rsum = 1.0 / (double)sum;  // the three IEEE quotients b[v] / sum (quot_f32)
    for (int v = 0; v < 3; ++v) b[v] = quot_f32(b[v], rsum);
    const float *Nm = sc.nmat + (size_t)sc.obj[t] * 9;
    float Tw[3];
    mat3_vec3(Nm, sc.tan + (size_t)t * 3, Tw);
    float Nv[3][3], Tv[3][3], Bv[3][3];
    for (int v = 0; v < 3; ++v) {
        mat3_vec3(Nm, sc.nrm + (size_t)t * 9 + v * 3, Nv[v]);
        // phong.vert:48-52 per vertex: T = normalMatrix * tangent, re-orthogonalized against N
        if (sc.vtan) mat3_vec3(Nm, sc.vtan + (size_t)t * 9 + v * 3, Tw);
        float dtn = dot3(Tw, Nv[v]);
        float tt[3] = {Tw[0] - dtn * Nv[v][0], Tw[1] - dtn * Nv[v][1], Tw[2] - dtn * Nv[v][2]};
        nrm3(tt, Tv[v]);
        cross3(Nv[v], Tv[v], Bv[v]);
    }
    float P[3], N[3], T[3], B[3];
    for (int k = 0; k < 3; ++k) {
        P[k] = interp3(b, Wp[0][k], Wp[1][k], Wp[2][k]);
        N[k] = interp3(b, Nv[0][k], Nv[1][k], Nv[2][k]);
        T[k] = interp3(b, Tv[0][k], Tv[1][k], Tv[2][k]);
        B[k] = interp3(b, Bv[0][k], Bv[1][k], Bv[2][k]);
    }
    uint32_t alb = sc.albedo[t];
    const int tid = sc.texid ? sc.texid[t] : -1;
    if (tid >= 0) {  // texture(diffuseMap, fragTexcoord) (phong.frag:449), stored as RGBA8
        const float *st = sc.uv + (size_t)t * 6;
        const float ts = interp3(b, st[0], st[2], st[4]), tt = interp3(b, st[1], st[3], st[5]);
        float rho2 = 0.0f;
        if (tid < sc.tex.count) {
            // The unnormalized barycentrics (before the division by sum) are
            // affine in NDC: d/dxn of b[i] is cy_j cw_k - cw_j cy_k and d/dyn
            // is cw_j cx_k - cx_j cw_k, (i, j, k) cyclic, so ds/dxn =
            // sum_i d b[i]/dxn (s_i - s) / sum exactly; one pixel is 2/W (2/H) in NDC.
            float gx[3], gy[3];
            for (int i = 0; i < 3; ++i) {
                const int j = (i + 1) % 3, k = (i + 2) % 3;
                gx[i] = clip[j][1] * clip[k][3] - clip[j][3] * clip[k][1];
                gy[i] = clip[j][3] * clip[k][0] - clip[j][0] * clip[k][3];
            }
            const int4 inf = sc.tex.info[tid];
            const float TW = (float)inf.y, TH = (float)inf.z;
            const float kx = 2.0f / (float)cm.W, ky = 2.0f / (float)cm.H;
            const float sx = (gx[0] * (st[0] - ts) + gx[1] * (st[2] - ts)) + gx[2] * (st[4] - ts);
            const float tx = (gx[0] * (st[1] - tt) + gx[1] * (st[3] - tt)) + gx[2] * (st[5] - tt);
            const float sy = (gy[0] * (st[0] - ts) + gy[1] * (st[2] - ts)) + gy[2] * (st[4] - ts);
            const float ty = (gy[0] * (st[1] - tt) + gy[1] * (st[3] - tt)) + gy[2] * (st[5] - tt);
            const float dudx = ((sx / sum) * kx) * TW, dvdx = ((tx / sum) * kx) * TH;
            const float dudy = ((sy / sum) * ky) * TW, dvdy = ((ty / sum) * ky) * TH;
            rho2 = fmaxf(dudx * dudx + dvdx * dvdx, dudy * dudy + dvdy * dvdy);
        }
        const float4 c = tex_sample(sc.tex, tid, ts, tt, rho2);
        alb = pack_unorm4x8(c.x, c.y, c.z, c.w);
    }
    const float rough = sc.rough ? sc.rough[sc.obj[t]] : 0.5f;
    o[0] = make_float4(P[0], P[1], P[2], __uint_as_float(alb));
    o[1] = make_float4(N[0], N[1], N[2], rough);
    o[2] = make_float4(T[0], T[1], T[2], __uint_as_float(1u));
    o[3] = make_float4(B[0], B[1], B[2], __uint_as_float((uint32_t)t));
}

hipError_t launch_gbuffer(const DevScene &sc, const vct_camera &cam, uint64_t *vis, vct_gbuffer_texel *gbuf,
                          RasterScratch &rs, hipStream_t s) {
    const size_t np = (size_t)cam.width * cam.height;
    hipError_t e = launch_fill_u64(vis, ~0ull, np, s);
    if (e != hipSuccess) return e;
    CamMats cm;
    // pv = projection * view, the matrix product phong.vert applies before model
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float acc = cam.projection[0 * 4 + r] * cam.view[c * 4 + 0];
            acc = acc + cam.projection[1 * 4 + r] * cam.view[c * 4 + 1];
            acc = acc + cam.projection[2 * 4 + r] * cam.view[c * 4 + 2];
            acc = acc + cam.projection[3 * 4 + r] * cam.view[c * 4 + 3];
            cm.pv[c * 4 + r] = acc;
        }
    cm.W = cam.width;
    cm.H = cam.height;
    cm.cull = cam.cull_back;
    if (sc.ntri > 0) {
        const size_t nslots = (size_t)sc.ntri * 2;
        hipLaunchKernelGGL(k_gbuf_setup, dim3((sc.ntri + 255) / 256), dim3(256), 0, s, sc, cm, rs.rec, rs.attr,
                           rs.counts, rs.bcounts);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = raster_scan(rs, nslots, true, s)) != hipSuccess) return e;
        GbufFrag f{rs.attr, (unsigned long long *)vis, cam.width};
        hipLaunchKernelGGL(k_raster_items<GbufFrag>, dim3(raster_grid<GbufFrag, false>()), dim3(256), 0, s, rs.rec, rs.offs,
                           rs.chunk, rs.chunk_cap, (int)nslots, f);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(k_block_slots, dim3(kRasterBlocks), dim3(256), 0, s, rs.boffs, rs.bchunk, rs.chunk_cap,
                           (int)nslots, rs.bslots, rs.bslots_cap);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        GbufBlockFrag fb{rs.attr, (unsigned long long *)vis, cam.width, cam.height, make_float4(0.f, 0.f, 0.f, 0.f)};
        hipLaunchKernelGGL(k_raster_blocks<GbufBlockFrag>, dim3(kRasterBlocks), dim3(256), 0, s, rs.rec, rs.boffs,
                           rs.bchunk, rs.chunk_cap, (int)nslots, rs.bslots, rs.bslots_cap, fb);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    dim3 grid((cam.width + 15) / 16, (cam.height + 15) / 16);
    hipLaunchKernelGGL(k_gbuf_resolve, grid, dim3(256), 0, s, sc, cm, (const unsigned long long *)vis, gbuf);
    return hipGetLastError();
}

}  // namespace vctk
