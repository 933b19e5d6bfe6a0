// context.hip -- the C ABI of include/vct.h: context lifetime, scene/light
// upload, pass dispatch on one HIP stream, timing, readback.
//
// The reference keeps its state in GL objects owned by VCT (voxel textures,
// src/Application.h:107-156), Application (shadow FBO, timers) and function
// statics (warpmap, fill-holes temporaries).  Here one vct_ctx owns every
// device buffer; all passes are enqueued on the context's stream, whose order
// replaces the reference's glMemoryBarrier calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <random>

#include "learned.h"
#include "vct_internal.h"

using namespace vctk;

enum PassId { P_SHADOW = 0, P_VOXELIZE, P_RADIANCE, P_FILL, P_MIP, P_GBUFFER, P_TRACE, P_LEARNED, P_COMPOSE, P_WARP, P_COUNT };

struct TimedPair {
    int pass;
    hipEvent_t a, b;
};

struct vct_ctx {
    vct_grid_desc desc{};
    hipStream_t stream = nullptr;
    // a second stream for work inside one pass that does not depend on the
    // rest of that pass (the albedo chain beside the radiance chain); forked
    // from and joined back to `stream` by events, so stream order holds
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::string err;
    // scene
    int ntri = 0, nobj = 0, has_rough = 0, have_transforms = 0;
    float *d_pos = nullptr, *d_nrm = nullptr, *d_tan = nullptr, *d_rough = nullptr;
    float *d_vtan = nullptr;  // per-corner tangents (vct_mesh.vertex_tangents), or null
    float *d_uv = nullptr;      // per-corner texture coordinates (vct_mesh.texcoords), or null
    int32_t *d_texid = nullptr;  // per-triangle diffuse map (vct_mesh.texture), or null
    uint32_t *d_tex_texels = nullptr;  // vct_set_textures: every level of every map
    int4 *d_tex_info = nullptr, *d_tex_levels = nullptr;
    int ntex = 0;
    uint32_t *d_albedo = nullptr, *d_obj = nullptr;
    float *d_model = nullptr, *d_nmat = nullptr;
    // pinned staging ring of the per-frame transform upload: slot k holds the
    // model and normal matrices of one vct_set_transforms call until the copy
    // out of it has run on the stream (xf_done[k]), so the host never waits for
    // the frame in flight
    static constexpr int kXfSlots = 3;
    float *h_xf[kXfSlots] = {};
    hipEvent_t xf_done[kXfSlots] = {};
    bool xf_recorded[kXfSlots] = {};
    int xf_next = 0;
    // lights
    std::vector<vct_light> lights;
    std::vector<float *> d_shadow;
    std::vector<float4 *> d_shadowq;  // footprint layout of each map (DevLight::shadowquad)
    std::vector<DevLight> hlights;
    DevLight *d_lights = nullptr;
    // voxel grids
    uint32_t *d_grid_albedo = nullptr, *d_grid_normal = nullptr;
    uint32_t *d_sums = nullptr;  // VCT_ACCUM_SUM: 8 words per voxel (allocated on first use)
    bool have_sums = false;      // the last vct_voxelize accumulated sums
    // the albedo (voxelColor) chain: levels 1..levels-1, RGBA8 linear x-fastest
    uint32_t *d_albedo_mips = nullptr;
    uint32_t *albedo_lv[VCT_MAX_LEVELS] = {};
    uint8_t *d_litmask = nullptr;
    uint32_t *d_partials = nullptr;  // per-workgroup statistics partials
    char *d_pyr = nullptr;
    uint8_t *d_zero0 = nullptr;  // PyrDesc::zero0
    float *d_scales = nullptr;
    size_t pyr_bytes = 0, scale_floats = 0;
    PyrDesc pyr{};
    LevelInfo *d_lv = nullptr;
    void *d_fill_tmp = nullptr;
    vct_stats *d_stats = nullptr;
    bool have_voxels = false, have_radiance = false, have_mips = false;
    // bit 1 of PyrDesc::zero0 describes the albedo grid as it is (set by the
    // injection, cleared by every other writer of the grid)
    bool albedo_zero_valid = false;
    // bit 2 describes both voxel grids as they are: set by an injection that
    // follows the standard voxelizer in the max or average mode (every voxel it
    // writes gets a nonzero albedo word, alpha 1 or a count), cleared by every
    // other writer of the grids
    bool grids_std_vox = false, grid_zero_valid = false;
    bool pyr_nonneg = true;  // level 0 was injected (not uploaded): no negative texels
    // raster scratch
    RasterScratch rs{};
    // camera images
    int gw = 0, gh = 0;
    vct_gbuffer_texel *d_gbuf = nullptr;
    uint64_t *d_vis = nullptr;
    float4 *d_ind = nullptr, *d_spec = nullptr, *d_color = nullptr;
    // what the last vct_trace of the current G-buffer left in d_ind / d_spec
    bool have_trace = false, trace_occluded = false, trace_spec = false;
    // the warp texture: 32^3 occupancy, per-axis prefix sums, the float4 texels
    uint32_t *d_warp_occ = nullptr;
    int4 *d_warp_part = nullptr;
    float4 *d_warpmap = nullptr;
    bool have_warpmap = false;
    bool have_gbuf = false;
    // staging for downloads
    float4 *d_stage = nullptr;
    size_t stage_bytes = 0;
    // learned gather
    vct_gbuffer_texel *d_points = nullptr;
    int npoints = 0;
    float4 *d_labels = nullptr;
    uint16_t *d_X = nullptr;
    bool have_prepared = false, have_weights = false;
    MlpBuffers mlp;
    float *d_loss = nullptr;
    int loss_cap = 0, last_train_blocks = 0;
    int adam_t = 0;  // Adam steps taken since the weights were initialised or uploaded
    uint32_t *d_tile_count = nullptr;  // fallback tiles of the last learned gather: count, then the tiles
    float4 *d_reference = nullptr;
    // timing
    bool timing = true;
    std::vector<hipEvent_t> ev_pool;
    std::vector<TimedPair> pending;
    vct_pass_times last{};
};

static thread_local std::string g_err;
static int ensure_mlp(vct_ctx *c);

static int fail(vct_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g_err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                                        \
    do {                                                                                                       \
        hipError_t _e = (expr);                                                                                \
        if (_e != hipSuccess) return fail((c), VCT_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

template <class T>
static hipError_t dalloc(T **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    return hipMalloc((void **)p, bytes);
}
template <class T>
static void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

static void free_xf_ring(vct_ctx *c) {
    for (int k = 0; k < vct_ctx::kXfSlots; ++k) {
        if (c->xf_done[k]) (void)hipEventDestroy(c->xf_done[k]);
        if (c->h_xf[k]) (void)hipHostFree(c->h_xf[k]);
        c->xf_done[k] = nullptr;
        c->h_xf[k] = nullptr;
        c->xf_recorded[k] = false;
    }
    c->xf_next = 0;
}

static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
static int ilog2(int v) {
    int l = 0;
    while ((1 << (l + 1)) <= v) ++l;
    return l;
}
static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// ---------------------------------------------------------------- timing
static hipEvent_t take_event(vct_ctx *c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct PassTimer {
    vct_ctx *c;
    int pass;
    hipEvent_t a = nullptr;
    PassTimer(vct_ctx *c_, int p) : c(c_), pass(p) {
        if (c->timing) {
            a = take_event(c);
            if (a) (void)hipEventRecord(a, c->stream);
        }
    }
    void stop() {
        if (!a) return;
        hipEvent_t b = take_event(c);
        if (b && hipEventRecord(b, c->stream) == hipSuccess) c->pending.push_back({pass, a, b});
        else c->ev_pool.push_back(a);
        a = nullptr;
    }
    ~PassTimer() { stop(); }
};

// ---------------------------------------------------------------- pyramid
static void free_grids(vct_ctx *c) {
    dfree(c->d_grid_albedo);
    dfree(c->d_grid_normal);
    dfree(c->d_sums);
    c->have_sums = false;
    dfree(c->d_albedo_mips);
    for (auto &p : c->albedo_lv) p = nullptr;
    dfree(c->d_litmask);
    dfree(c->d_pyr);
    dfree(c->d_zero0);
    dfree(c->d_scales);
    dfree(c->d_lv);
    dfree(c->d_fill_tmp);
    c->have_voxels = c->have_radiance = c->have_mips = false;
}

static int alloc_grids(vct_ctx *c) {
    const int dim = c->desc.dim;
    const size_t nv = (size_t)dim * dim * dim;
    HIPCHK(c, dalloc(&c->d_grid_albedo, nv * 4));
    HIPCHK(c, dalloc(&c->d_grid_normal, nv * 4));
    HIPCHK(c, dalloc(&c->d_litmask, nv));
    c->albedo_zero_valid = false;
    c->grids_std_vox = c->grid_zero_valid = false;
    HIPCHK(c, hipMemsetAsync(c->d_grid_albedo, 0, nv * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_grid_normal, 0, nv * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_litmask, 0, nv, c->stream));
    {
        size_t words = 0;
        for (int l = 1; l < c->desc.levels; ++l) {
            const size_t sz = (size_t)std::max(1, dim >> l);
            words += sz * sz * sz;
        }
        if (words) {
            HIPCHK(c, dalloc(&c->d_albedo_mips, words * 4));
            HIPCHK(c, hipMemsetAsync(c->d_albedo_mips, 0, words * 4, c->stream));
            size_t off = 0;
            for (int l = 1; l < c->desc.levels; ++l) {
                const size_t sz = (size_t)std::max(1, dim >> l);
                c->albedo_lv[l] = c->d_albedo_mips + off;
                off += sz * sz * sz;
            }
        }
    }
    PyrDesc &P = c->pyr;
    memset(&P, 0, sizeof(P));
    P.dim = dim;
    P.nlevels = c->desc.levels;
    P.storage = c->desc.storage;
    P.aniso = c->desc.anisotropic;
    // one allocation: per level, the faces' bricks interleaved (face f of brick
    // b at element 64 * (b * nfaces + f)), then the fp16-pair stage copy of
    // fp8 levels; every offset 256-byte aligned
    size_t off = 0, soff = 0;
    for (int l = 0; l < P.nlevels; ++l) {
        LevelInfo &L = P.lv[l];
        L.size = std::max(1, dim >> l);
        L.nb = (L.size + 3) / 4;
        L.fmt = P.storage == VCT_STORE_FP32 ? 1 : ((P.storage == VCT_STORE_FP8 && l >= 3) ? 2 : 0);
        L.nfaces = (P.aniso && l >= 1) ? 6 : 1;
        const uint32_t nbp = (uint32_t)L.nb + 2;  // a border brick on each side
        L.sx = 64u * (uint32_t)L.nfaces;
        L.sy = L.sx * nbp;
        L.sz = L.sy * nbp;
        L.pad = L.sx + L.sy + L.sz;
        const size_t nelem = (size_t)L.sz * nbp;
        if (nelem > 0xffffffffull)
            return fail(c, VCT_ERR_INVALID, "radiance level %d of %zu elements exceeds the 32-bit range", l, nelem);
        L.nelem = (uint32_t)nelem;
        const size_t esz = L.fmt == 2 ? 4 : 16;
        if (off + align256(nelem * esz) > 0xffffff00ull)
            return fail(c, VCT_ERR_INVALID, "radiance pyramid exceeds the 32-bit offset range at level %d", l);
        L.base = (uint32_t)(off + (size_t)L.pad * esz);
        off += align256(nelem * esz);
        if (L.fmt == 2) {
            if (off + align256(nelem * 16) > 0xffffff00ull)
                return fail(c, VCT_ERR_INVALID, "radiance pyramid exceeds the 32-bit offset range at level %d", l);
            L.stage = (uint32_t)(off + (size_t)L.pad * 16);
            off += align256(nelem * 16);
            // scales[scale + element / 64] for elements relative to base: the
            // interior bricks (the border holds zero codes, read by nobody)
            L.scale = (uint32_t)soff;
            soff += nelem / 64;
        } else {
            L.stage = L.base;
            L.scale = 0;
        }
        if (off > 0xffffff00ull)
            return fail(c, VCT_ERR_INVALID, "radiance pyramid of %zu bytes exceeds the 32-bit offset range", off);
    }
    c->pyr_bytes = off;
    c->scale_floats = soff;
    HIPCHK(c, dalloc(&c->d_pyr, off));
    HIPCHK(c, hipMemsetAsync(c->d_pyr, 0, off, c->stream));  // the border bricks and halves stay 0
    if (soff) {
        HIPCHK(c, dalloc(&c->d_scales, soff * 4));
        HIPCHK(c, hipMemsetAsync(c->d_scales, 0, soff * 4, c->stream));
    }
    {
        const size_t nb0 = (size_t)P.lv[0].nb;
        const size_t zbytes = (nb0 * nb0 * nb0 + 3) / 4 * 4;
        HIPCHK(c, dalloc(&c->d_zero0, zbytes));
        HIPCHK(c, hipMemsetAsync(c->d_zero0, 0, zbytes, c->stream));  // unknown until the first injection
    }
    P.pyr = c->d_pyr;
    P.zero0 = c->d_zero0;
    P.bytes = (uint32_t)off;
    P.scales = c->d_scales;
    HIPCHK(c, dalloc(&c->d_lv, sizeof(LevelInfo) * P.nlevels));
    HIPCHK(c, hipMemcpyAsync(c->d_lv, P.lv, sizeof(LevelInfo) * P.nlevels, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    P.d_lv = c->d_lv;
    return VCT_OK;
}

static int validate_desc(const vct_grid_desc *d, vct_grid_desc *out, std::string *warn) {
    if (!d) return fail(nullptr, VCT_ERR_INVALID, "vct_create: desc is NULL");
    if (!is_pow2(d->dim) || d->dim < 8 || d->dim > 512)
        return fail(nullptr, VCT_ERR_INVALID, "voxel dim %d must be a power of two in [8, 512]", d->dim);
    if (d->storage < VCT_STORE_FP16 || d->storage > VCT_STORE_FP8)
        return fail(nullptr, VCT_ERR_INVALID, "unknown storage format %d", d->storage);
    for (int k = 0; k < 3; ++k)
        if (!(d->vmax[k] > d->vmin[k]))
            return fail(nullptr, VCT_ERR_INVALID, "grid extent max[%d] must exceed min[%d]", k, k);
    *out = *d;
    // VCT::remake: voxelLevels = clamp(levels, 0, log2(dim) + 1) with a warning
    const int maxl = std::min(ilog2(d->dim) + 1, VCT_MAX_LEVELS);
    int lv = d->levels < 1 ? 1 : (d->levels > maxl ? maxl : d->levels);
    if (lv != d->levels && warn) {
        char b[160];
        snprintf(b, sizeof(b), "Attempted remaking VCT with invalid number of levels, clamped %d to %d", d->levels,
                 lv);
        *warn = b;
    }
    out->levels = lv;
    out->anisotropic = d->anisotropic ? 1 : 0;
    return VCT_OK;
}

// ---------------------------------------------------------------- ABI
extern "C" {

int vct_abi_version(void) { return VCT_ABI_VERSION; }

const char *vct_last_error(const vct_ctx *c) { return c ? c->err.c_str() : g_err.c_str(); }

int vct_create(vct_ctx **out, const vct_grid_desc *desc) {
    if (!out) return fail(nullptr, VCT_ERR_INVALID, "vct_create: out is NULL");
    *out = nullptr;
    vct_grid_desc d;
    std::string warn;
    int rc = validate_desc(desc, &d, &warn);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, VCT_ERR_NODEVICE, "no HIP device available");
    if (d.device < 0 || d.device >= ndev)
        return fail(nullptr, VCT_ERR_INVALID, "device %d out of range (have %d)", d.device, ndev);
    if (hipSetDevice(d.device) != hipSuccess) return fail(nullptr, VCT_ERR_NODEVICE, "hipSetDevice failed");
    vct_ctx *c = new vct_ctx();
    c->desc = d;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(nullptr, VCT_ERR_HIP, "hipStreamCreate failed");
    }
    if (hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess) {
        vct_destroy(c);
        return fail(nullptr, VCT_ERR_HIP, "hipStreamCreate (aux) failed");
    }
    if (hipMalloc((void **)&c->d_partials, kStatPartials * 4 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&c->d_stats, sizeof(vct_stats)) != hipSuccess ||
        hipMemset(c->d_stats, 0, sizeof(vct_stats)) != hipSuccess) {
        vct_destroy(c);
        return fail(nullptr, VCT_ERR_NOMEM, "stats allocation failed");
    }
    rc = alloc_grids(c);
    if (rc) {
        g_err = c->err;
        vct_destroy(c);
        return rc;
    }
    // this device's address of the static scratch of vct_normalize_inject_noalloc
    // (looked up once per device, with the device current)
    if (noalloc_scratch_setup(d.device) != hipSuccess) {
        vct_destroy(c);
        return fail(nullptr, VCT_ERR_HIP, "hipGetSymbolAddress of the no-alloc scratch failed");
    }
    c->err = warn;
    *out = c;
    return VCT_OK;
}

void vct_destroy(vct_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->aux) (void)hipStreamSynchronize(c->aux);
    free_grids(c);
    dfree(c->d_pos);
    dfree(c->d_nrm);
    dfree(c->d_tan);
    dfree(c->d_rough);
    dfree(c->d_vtan);
    dfree(c->d_uv);
    dfree(c->d_texid);
    dfree(c->d_tex_texels);
    dfree(c->d_tex_info);
    dfree(c->d_tex_levels);
    dfree(c->d_albedo);
    dfree(c->d_obj);
    dfree(c->d_model);
    dfree(c->d_nmat);
    free_xf_ring(c);
    for (auto &p : c->d_shadow) dfree(p);
    for (auto &p : c->d_shadowq) dfree(p);
    dfree(c->d_lights);
    dfree(c->rs.rec);
    dfree(c->rs.attr);
    dfree(c->rs.counts);
    dfree(c->rs.offs);
    dfree(c->rs.chunk);
    dfree(c->rs.bcounts);
    dfree(c->rs.boffs);
    dfree(c->rs.bchunk);
    dfree(c->rs.bslots);
    dfree(c->rs.tsum);
    dfree(c->d_gbuf);
    dfree(c->d_vis);
    dfree(c->d_ind);
    dfree(c->d_spec);
    dfree(c->d_color);
    dfree(c->d_warp_occ);
    dfree(c->d_warp_part);
    dfree(c->d_warpmap);
    dfree(c->d_stage);
    dfree(c->d_points);
    dfree(c->d_labels);
    dfree(c->d_X);
    dfree(c->mlp.master);
    dfree(c->mlp.m1);
    dfree(c->mlp.m2);
    dfree(c->mlp.w1);
    dfree(c->mlp.w2);
    dfree(c->mlp.w3);
    dfree(c->mlp.w4);
    dfree(c->mlp.w2t);
    dfree(c->mlp.w3t);
    dfree(c->mlp.w4t);
    dfree(c->mlp.partials);
    dfree(c->d_loss);
    dfree(c->d_tile_count);
    dfree(c->d_reference);
    dfree(c->d_stats);
    dfree(c->d_partials);
    for (auto &tp : c->pending) {
        (void)hipEventDestroy(tp.a);
        (void)hipEventDestroy(tp.b);
    }
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->aux) (void)hipStreamSynchronize(c->aux);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->aux) (void)hipStreamDestroy(c->aux);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int vct_remake(vct_ctx *c, int dim, int levels) {
    if (!c) return fail(nullptr, VCT_ERR_INVALID, "ctx is NULL");
    vct_grid_desc d = c->desc, nd;
    d.dim = dim;
    d.levels = levels;
    std::string warn;
    int rc = validate_desc(&d, &nd, &warn);
    if (rc) {
        c->err = g_err;
        return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipStreamSynchronize(c->aux));
    free_grids(c);
    c->desc = nd;
    rc = alloc_grids(c);
    if (rc) return rc;
    c->err = warn;
    return VCT_OK;
}

int vct_get_grid(const vct_ctx *c, vct_grid_desc *out) {
    if (!c || !out) return VCT_ERR_INVALID;
    *out = c->desc;
    return VCT_OK;
}

int vct_set_extent(vct_ctx *c, const float center[3], const float vmin[3], const float vmax[3]) {
    if (!c || !center || !vmin || !vmax) return fail(c, VCT_ERR_INVALID, "vct_set_extent: NULL argument");
    for (int k = 0; k < 3; ++k)
        if (!(vmax[k] > vmin[k])) return fail(c, VCT_ERR_INVALID, "extent max[%d] must exceed min[%d]", k, k);
    for (int k = 0; k < 3; ++k) {
        c->desc.center[k] = center[k];
        c->desc.vmin[k] = vmin[k];
        c->desc.vmax[k] = vmax[k];
    }
    return VCT_OK;
}

static int ensure_raster(vct_ctx *c, size_t slots) {
    if (slots <= c->rs.cap_slots) return VCT_OK;
    dfree(c->rs.rec);
    dfree(c->rs.attr);
    dfree(c->rs.counts);
    dfree(c->rs.offs);
    dfree(c->rs.bcounts);
    dfree(c->rs.boffs);
    dfree(c->rs.tsum);
    HIPCHK(c, dalloc(&c->rs.rec, slots * 4 * sizeof(int4)));
    HIPCHK(c, dalloc(&c->rs.attr, slots * 8 * sizeof(float4)));
    HIPCHK(c, dalloc(&c->rs.counts, (slots + 1) * sizeof(uint32_t)));
    HIPCHK(c, dalloc(&c->rs.offs, (slots + 1) * sizeof(uint64_t)));
    HIPCHK(c, dalloc(&c->rs.bcounts, (slots + 1) * sizeof(uint32_t)));
    HIPCHK(c, dalloc(&c->rs.boffs, (slots + 1) * sizeof(uint64_t)));
    c->rs.scan_tiles_cap = raster_scan_tiles(slots);
    HIPCHK(c, dalloc(&c->rs.tsum, (4 * c->rs.scan_tiles_cap + 4) * sizeof(uint64_t)));
    if (!c->rs.chunk) {
        c->rs.chunk_cap = (size_t)1 << 22;  // chunk starts for 256M items; beyond that a full search is used
        HIPCHK(c, dalloc(&c->rs.chunk, c->rs.chunk_cap * sizeof(uint32_t)));
        HIPCHK(c, dalloc(&c->rs.bchunk, c->rs.chunk_cap * sizeof(uint32_t)));
        c->rs.bslots_cap = (size_t)1 << 24;  // 64 MiB: blocks past this use the walk's own search
        HIPCHK(c, dalloc(&c->rs.bslots, c->rs.bslots_cap * sizeof(uint32_t)));
    }
    c->rs.cap_slots = slots;
    return VCT_OK;
}

int vct_set_scene(vct_ctx *c, const vct_mesh *m) {
    if (!c || !m) return fail(c, VCT_ERR_INVALID, "vct_set_scene: NULL argument");
    if (m->num_tris < 0 || m->num_objects <= 0)
        return fail(c, VCT_ERR_INVALID, "vct_set_scene: num_tris %d / num_objects %d invalid", m->num_tris,
                    m->num_objects);
    if (m->num_tris > (1 << 28)) return fail(c, VCT_ERR_INVALID, "vct_set_scene: too many triangles");
    if (m->num_tris > 0 && (!m->positions || !m->normals || !m->tangents || !m->albedo || !m->object))
        return fail(c, VCT_ERR_INVALID, "vct_set_scene: NULL geometry array");
    if (m->num_tris > 0 && m->texture && !m->texcoords)
        return fail(c, VCT_ERR_INVALID, "vct_set_scene: texture indices without texcoords");
    for (int t = 0; t < m->num_tris; ++t)
        if (m->object[t] >= (uint32_t)m->num_objects)
            return fail(c, VCT_ERR_INVALID, "vct_set_scene: triangle %d references object %u >= %d", t,
                        m->object[t], m->num_objects);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c->d_pos);
    dfree(c->d_nrm);
    dfree(c->d_tan);
    dfree(c->d_albedo);
    dfree(c->d_obj);
    dfree(c->d_rough);
    dfree(c->d_vtan);
    dfree(c->d_uv);
    dfree(c->d_texid);
    dfree(c->d_model);
    dfree(c->d_nmat);
    free_xf_ring(c);  // the stream is idle: no copy reads the ring
    const size_t T = (size_t)m->num_tris;
    HIPCHK(c, dalloc(&c->d_pos, T * 9 * 4));
    HIPCHK(c, dalloc(&c->d_nrm, T * 9 * 4));
    HIPCHK(c, dalloc(&c->d_tan, T * 3 * 4));
    HIPCHK(c, dalloc(&c->d_albedo, T * 4));
    HIPCHK(c, dalloc(&c->d_obj, T * 4));
    HIPCHK(c, dalloc(&c->d_model, (size_t)m->num_objects * 16 * 4));
    HIPCHK(c, dalloc(&c->d_nmat, (size_t)m->num_objects * 9 * 4));
    for (int k = 0; k < vct_ctx::kXfSlots; ++k) {
        HIPCHK(c, hipHostMalloc((void **)&c->h_xf[k], (size_t)m->num_objects * 25 * 4, hipHostMallocDefault));
        HIPCHK(c, hipEventCreateWithFlags(&c->xf_done[k], hipEventDisableTiming));
    }
    if (T) {
        HIPCHK(c, hipMemcpy(c->d_pos, m->positions, T * 9 * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_nrm, m->normals, T * 9 * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_tan, m->tangents, T * 3 * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_albedo, m->albedo, T * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_obj, m->object, T * 4, hipMemcpyHostToDevice));
    }
    if (T && m->vertex_tangents) {
        HIPCHK(c, dalloc(&c->d_vtan, T * 9 * 4));
        HIPCHK(c, hipMemcpy(c->d_vtan, m->vertex_tangents, T * 9 * 4, hipMemcpyHostToDevice));
    }
    if (T && m->texture && m->texcoords) {  // the diffuse maps' indices and coordinates
        HIPCHK(c, dalloc(&c->d_texid, T * 4));
        HIPCHK(c, dalloc(&c->d_uv, T * 6 * 4));
        HIPCHK(c, hipMemcpy(c->d_texid, m->texture, T * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_uv, m->texcoords, T * 6 * 4, hipMemcpyHostToDevice));
    }
    c->has_rough = m->roughness != nullptr;
    if (c->has_rough) {
        HIPCHK(c, dalloc(&c->d_rough, (size_t)m->num_objects * 4));
        HIPCHK(c, hipMemcpy(c->d_rough, m->roughness, (size_t)m->num_objects * 4, hipMemcpyHostToDevice));
    }
    c->ntri = m->num_tris;
    c->nobj = m->num_objects;
    c->have_transforms = 0;
    int rc = ensure_raster(c, std::max<size_t>(T * 2, 64));
    if (rc) return rc;
    // identity transforms until the caller sets its own
    std::vector<float> id((size_t)c->nobj * 16, 0.0f);
    for (int o = 0; o < c->nobj; ++o) id[(size_t)o * 16 + 0] = id[(size_t)o * 16 + 5] = id[(size_t)o * 16 + 10] =
        id[(size_t)o * 16 + 15] = 1.0f;
    return vct_set_transforms(c, id.data(), c->nobj);
}

int vct_set_textures(vct_ctx *c, const vct_texture *tex, int n) {
    if (!c) return fail(nullptr, VCT_ERR_INVALID, "ctx is NULL");
    if (n < 0 || (n > 0 && !tex)) return fail(c, VCT_ERR_INVALID, "vct_set_textures: %d textures at %p", n, (const void *)tex);
    std::vector<int4> info((size_t)n), levels;
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        const vct_texture &t = tex[i];
        if (t.width < 1 || t.height < 1 || t.width > 32768 || t.height > 32768)
            return fail(c, VCT_ERR_INVALID, "vct_set_textures: texture %d is %d x %d", i, t.width, t.height);
        int q = 0;
        while ((std::max(t.width, t.height) >> (q + 1)) > 0) ++q;
        if (t.levels < 0 || t.levels > q + 1 || (t.levels > 0 && !t.texels))
            return fail(c, VCT_ERR_INVALID, "vct_set_textures: texture %d has %d levels (at most %d)", i, t.levels, q + 1);
        info[i] = make_int4((int)levels.size(), t.width, t.height, (t.levels == q + 1 ? 1 : 0) | (q << 8));
        for (int l = 0; l < t.levels; ++l) {
            const int w = std::max(1, t.width >> l), h = std::max(1, t.height >> l);
            levels.push_back(make_int4((int)(uint32_t)total, w, h, 0));
            total += (size_t)w * h;
        }
        if (total >= ((size_t)1 << 32)) return fail(c, VCT_ERR_INVALID, "vct_set_textures: more than 2^32 texels");
        if (t.levels == 0) levels.push_back(make_int4(0, t.width, t.height, 0));  // never read: incomplete
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c->d_tex_texels);
    dfree(c->d_tex_info);
    dfree(c->d_tex_levels);
    c->ntex = 0;
    if (n == 0) return VCT_OK;
    HIPCHK(c, dalloc(&c->d_tex_texels, std::max<size_t>(total, 1) * 4));
    HIPCHK(c, dalloc(&c->d_tex_info, info.size() * sizeof(int4)));
    HIPCHK(c, dalloc(&c->d_tex_levels, levels.size() * sizeof(int4)));
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        size_t cnt = 0;
        for (int l = 0; l < tex[i].levels; ++l)
            cnt += (size_t)std::max(1, tex[i].width >> l) * (size_t)std::max(1, tex[i].height >> l);
        if (cnt) HIPCHK(c, hipMemcpy(c->d_tex_texels + off, tex[i].texels, cnt * 4, hipMemcpyHostToDevice));
        off += cnt;
    }
    HIPCHK(c, hipMemcpy(c->d_tex_info, info.data(), info.size() * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_tex_levels, levels.data(), levels.size() * sizeof(int4), hipMemcpyHostToDevice));
    c->ntex = n;
    return VCT_OK;
}

// mat3(transpose(inverse(model))) as cofactor(A) / det(A) of the upper-left
// 3x3 block (voxelize.vert:18, phong.vert:37); column-major out.
static void normal_matrix(const float *m, float *out) {
#define A(r, cc) m[(cc)*4 + (r)]
    float c00 = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
    float c01 = A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2);
    float c02 = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
    float c10 = A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2);
    float c11 = A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0);
    float c12 = A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1);
    float c20 = A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1);
    float c21 = A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2);
    float c22 = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
    float det = (A(0, 0) * c00 + A(0, 1) * c01) + A(0, 2) * c02;
#undef A
    out[0] = c00 / det;
    out[1] = c10 / det;
    out[2] = c20 / det;
    out[3] = c01 / det;
    out[4] = c11 / det;
    out[5] = c21 / det;
    out[6] = c02 / det;
    out[7] = c12 / det;
    out[8] = c22 / det;
}

int vct_set_transforms(vct_ctx *c, const float *model, int n) {
    if (!c || !model) return fail(c, VCT_ERR_INVALID, "vct_set_transforms: NULL argument");
    if (n != c->nobj) return fail(c, VCT_ERR_INVALID, "vct_set_transforms: %d matrices for %d objects", n, c->nobj);
    // the host arrays are borrowed for the call only: they are copied into a
    // pinned ring slot whose previous copy has completed, and the device copy
    // runs from there in stream order (after the passes of earlier frames)
    const int k = c->xf_next;
    c->xf_next = (k + 1) % vct_ctx::kXfSlots;
    if (c->xf_recorded[k]) HIPCHK(c, hipEventSynchronize(c->xf_done[k]));
    float *h = c->h_xf[k];
    memcpy(h, model, (size_t)n * 64);
    for (int o = 0; o < n; ++o) normal_matrix(model + (size_t)o * 16, h + (size_t)n * 16 + (size_t)o * 9);
    HIPCHK(c, hipMemcpyAsync(c->d_model, h, (size_t)n * 64, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_nmat, h + (size_t)n * 16, (size_t)n * 36, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipEventRecord(c->xf_done[k], c->stream));
    c->xf_recorded[k] = true;
    c->have_transforms = 1;
    return VCT_OK;
}

int vct_set_lights(vct_ctx *c, const vct_light *lights, int n) {
    if (!c || (n > 0 && !lights)) return fail(c, VCT_ERR_INVALID, "vct_set_lights: NULL argument");
    if (n < 0 || n > kMaxLights) return fail(c, VCT_ERR_INVALID, "vct_set_lights: %d lights (max %d)", n, kMaxLights);
    for (int i = 0; i < n; ++i) {
        if (lights[i].type > 1u) return fail(c, VCT_ERR_INVALID, "light %d: unknown type %u", i, lights[i].type);
        if (lights[i].shadow_caster && (lights[i].shadow_size < 16 || lights[i].shadow_size > 8192))
            return fail(c, VCT_ERR_INVALID, "light %d: shadow_size %u out of [16, 8192]", i, lights[i].shadow_size);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // keep shadow maps whose size is unchanged (rebuilt per frame by vct_shadowmap(s))
    std::vector<float *> maps(n, nullptr);
    std::vector<float4 *> quads(n, nullptr);
    for (int i = 0; i < n; ++i) {
        const bool want = lights[i].shadow_caster != 0;
        if (want && i < (int)c->d_shadow.size() && c->d_shadow[i] && i < (int)c->lights.size() &&
            c->lights[i].shadow_caster && c->lights[i].shadow_size == lights[i].shadow_size) {
            maps[i] = c->d_shadow[i];
            quads[i] = c->d_shadowq[i];
            c->d_shadow[i] = nullptr;
            c->d_shadowq[i] = nullptr;
        } else if (want) {
            const size_t N = lights[i].shadow_size;
            HIPCHK(c, dalloc(&maps[i], N * N * 4));
            HIPCHK(c, dalloc(&quads[i], (N + 1) * (N + 1) * 16));
            if (launch_fill_u32((uint32_t *)maps[i], 0x3f800000u, N * N, c->stream) != hipSuccess ||
                launch_shadow_quads(maps[i], (int)N, quads[i], c->stream) != hipSuccess)
                return fail(c, VCT_ERR_HIP, "shadow map clear failed");
        }
    }
    for (auto &p : c->d_shadow) dfree(p);
    for (auto &p : c->d_shadowq) dfree(p);
    c->d_shadow = maps;
    c->d_shadowq = quads;
    c->lights.assign(lights, lights + n);
    c->hlights.resize(n);
    for (int i = 0; i < n; ++i) {
        DevLight &d = c->hlights[i];
        const vct_light &L = lights[i];
        memcpy(d.position, L.position, 12);
        d.range = L.range;
        memcpy(d.direction, L.direction, 12);
        d.intensity = L.intensity;
        memcpy(d.color, L.color, 12);
        d.type = L.type;
        d.enabled = L.enabled;
        d.shadow_caster = L.shadow_caster;
        d.shadow_size = L.shadow_size;
        d.perspective = L.perspective;
        memcpy(d.ls, L.ls, 64);
        memcpy(d.ls_inverse, L.ls_inverse, 64);
        d.shadowmap = maps[i];
        d.shadowquad = quads[i];
    }
    dfree(c->d_lights);
    if (n) {
        HIPCHK(c, dalloc(&c->d_lights, sizeof(DevLight) * n));
        HIPCHK(c, hipMemcpyAsync(c->d_lights, c->hlights.data(), sizeof(DevLight) * n, hipMemcpyHostToDevice,
                                 c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VCT_OK;
}

static DevScene dev_scene(const vct_ctx *c) {
    DevScene s;
    s.ntri = c->ntri;
    s.nobj = c->nobj;
    s.pos = c->d_pos;
    s.nrm = c->d_nrm;
    s.tan = c->d_tan;
    s.albedo = c->d_albedo;
    s.obj = c->d_obj;
    s.rough = c->has_rough ? c->d_rough : nullptr;
    s.vtan = c->d_vtan;
    s.uv = c->d_texid ? c->d_uv : nullptr;
    s.texid = c->d_texid;
    s.tex.texels = c->d_tex_texels;
    s.tex.info = c->d_tex_info;
    s.tex.levels = c->d_tex_levels;
    s.tex.count = c->ntex;
    s.model = c->d_model;
    s.nmat = c->d_nmat;
    return s;
}

#define LAUNCHCHK(c, expr)                                                                               \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) return fail((c), VCT_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e));   \
    } while (0)

int vct_shadowmap(vct_ctx *c, int li) {
    if (!c) return fail(nullptr, VCT_ERR_INVALID, "ctx is NULL");
    if (li < 0 || li >= (int)c->lights.size()) return fail(c, VCT_ERR_INVALID, "light index %d out of range", li);
    if (!c->lights[li].shadow_caster) return fail(c, VCT_ERR_INVALID, "light %d casts no shadow", li);
    if (!c->d_model) return fail(c, VCT_ERR_STATE, "vct_shadowmap: no scene");
    PassTimer pt(c, P_SHADOW);
    LAUNCHCHK(c, launch_shadowmaps(dev_scene(c), c->hlights.data(), c->d_lights, &li, 1, c->rs, c->stream));
    return VCT_OK;
}

int vct_shadowmaps(vct_ctx *c) {
    if (!c) return fail(nullptr, VCT_ERR_INVALID, "ctx is NULL");
    if (!c->d_model) return fail(c, VCT_ERR_STATE, "vct_shadowmaps: no scene");
    std::vector<int> li;
    for (int i = 0; i < (int)c->lights.size(); ++i)
        if (c->lights[i].shadow_caster) li.push_back(i);
    if (li.empty()) return VCT_OK;
    const size_t walk = std::min<size_t>(li.size(), (size_t)shadow_batch(c->ntri));
    int rc = ensure_raster(c, std::max<size_t>(walk * (size_t)c->ntri, 64));
    if (rc) return rc;
    PassTimer pt(c, P_SHADOW);
    LAUNCHCHK(c, launch_shadowmaps(dev_scene(c), c->hlights.data(), c->d_lights, li.data(), (int)li.size(), c->rs,
                                   c->stream));
    return VCT_OK;
}

int vct_voxelize(vct_ctx *c, const vct_voxelize_params *p) {
    if (!c || !p) return fail(c, VCT_ERR_INVALID, "vct_voxelize: NULL argument");
    if (p->atomic_max < VCT_ACCUM_AVG || p->atomic_max > VCT_ACCUM_SUM)
        return fail(c, VCT_ERR_INVALID, "vct_voxelize: atomic_max %d is not a VCT_ACCUM_* mode", p->atomic_max);
    if (p->axis_override < -1 || p->axis_override > 2)
        return fail(c, VCT_ERR_INVALID, "axis_override %d out of [-1, 2]", p->axis_override);
    if (!c->d_model) return fail(c, VCT_ERR_STATE, "vct_voxelize: no scene");
    if (p->warp_texture && !c->have_warpmap) return fail(c, VCT_ERR_STATE, "vct_voxelize: warp_texture without vct_warpmap");
    if (p->atomic_max == VCT_ACCUM_SUM && !c->d_sums) {
        const size_t nv = (size_t)c->desc.dim * c->desc.dim * c->desc.dim;
        HIPCHK(c, dalloc(&c->d_sums, nv * 8 * sizeof(uint32_t)));
    }
    PassTimer pt(c, P_VOXELIZE);
    const uint8_t *zg = c->grid_zero_valid ? c->pyr.zero0 : nullptr;
    c->grids_std_vox = c->grid_zero_valid = false;
    // the grids' clear runs on the second stream beside the triangle setup and
    // the scan, from a fork after everything enqueued so far; the fragment
    // walk waits for it
    HIPCHK(c, hipEventRecord(c->fork, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->aux, c->fork, 0));
    LAUNCHCHK(c, launch_clear_grids(c->d_grid_albedo, c->d_grid_normal, c->desc.dim, zg, c->aux));
    HIPCHK(c, hipEventRecord(c->join, c->aux));
    LAUNCHCHK(c, launch_voxelize(dev_scene(c), c->d_lights, (int)c->lights.size(), *p, c->desc.dim,
                                 c->d_grid_albedo, c->d_grid_normal, c->join, c->d_sums, c->d_stats, c->d_partials,
                                 p->warp_texture ? c->d_warpmap : nullptr, c->rs, c->stream));
    c->grids_std_vox = p->atomic_max != VCT_ACCUM_SUM;
    c->have_voxels = true;
    c->albedo_zero_valid = false;
    c->have_mips = false;
    c->have_sums = p->atomic_max == VCT_ACCUM_SUM;
    return VCT_OK;
}

int vct_voxelize_tessellated(vct_ctx *c, const vct_voxelize_params *p) {
    if (!c || !p) return fail(c, VCT_ERR_INVALID, "vct_voxelize_tessellated: NULL argument");
    if (p->atomic_max != VCT_ACCUM_AVG && p->atomic_max != VCT_ACCUM_MAX)
        return fail(c, VCT_ERR_INVALID, "vct_voxelize_tessellated: atomic_max %d is not VCT_ACCUM_AVG or _MAX",
                    p->atomic_max);
    if (!c->d_model) return fail(c, VCT_ERR_STATE, "vct_voxelize_tessellated: no scene");
    PassTimer pt(c, P_VOXELIZE);
    const uint8_t *zg = c->grid_zero_valid ? c->pyr.zero0 : nullptr;
    c->grids_std_vox = c->grid_zero_valid = false;  // its albedo alpha is the map's: a written voxel may read 0
    LAUNCHCHK(c, launch_voxelize_tess(dev_scene(c), c->desc.center, c->desc.vmin, c->desc.vmax, c->desc.dim,
                                      p->atomic_max, c->d_grid_albedo, c->d_grid_normal, zg, c->d_stats, c->rs,
                                      c->stream));
    c->have_voxels = true;
    c->albedo_zero_valid = false;
    c->have_mips = false;
    c->have_sums = false;
    return VCT_OK;
}

int vct_normalize_voxels(vct_ctx *c, float opacity) {
    if (!c) return fail(nullptr, VCT_ERR_INVALID, "ctx is NULL");
    if (!(opacity >= 0.0f && opacity <= 1.0f))
        return fail(c, VCT_ERR_INVALID, "voxel_set_opacity %f out of [0, 1]", (double)opacity);
    if (!c->have_sums)
        return fail(c, VCT_ERR_STATE, "vct_normalize_voxels: the last vct_voxelize did not accumulate sums");
    PassTimer pt(c, P_RADIANCE);
    c->albedo_zero_valid = false;
    c->grids_std_vox = c->grid_zero_valid = false;
    LAUNCHCHK(c, launch_normalize_voxels(c->d_sums, c->d_grid_albedo, c->d_grid_normal, c->desc.dim, opacity, c->pyr,
                                         c->d_stats, c->d_partials, c->stream));
    c->have_radiance = true;
    c->have_mips = false;
    c->pyr_nonneg = true;
    return VCT_OK;
}

int vct_warpmap(vct_ctx *c, const vct_voxelize_params *vox, const vct_warpmap_params *p) {
    if (!c || !vox || !p) return fail(c, VCT_ERR_INVALID, "vct_warpmap: NULL argument");
    if (vox->axis_override < -1 || vox->axis_override > 2)
        return fail(c, VCT_ERR_INVALID, "axis_override %d out of [-1, 2]", vox->axis_override);
    if (!(p->high_resolution > 0.0f) || !(p->low_resolution > 0.0f))
        return fail(c, VCT_ERR_INVALID, "vct_warpmap: resolutions must be positive");
    if (!c->d_model) return fail(c, VCT_ERR_STATE, "vct_warpmap: no scene");
    const size_t n = (size_t)kWarpDim * kWarpDim * kWarpDim;
    if (!c->d_warpmap) {
        HIPCHK(c, dalloc(&c->d_warp_occ, n * 4));
        HIPCHK(c, dalloc(&c->d_warp_part, n * sizeof(int4)));
        HIPCHK(c, dalloc(&c->d_warpmap, n * sizeof(float4)));
    }
    PassTimer pt(c, P_WARP);
    LAUNCHCHK(c, launch_warp_occupancy(dev_scene(c), *vox, c->d_warp_occ, c->rs, c->stream));
    LAUNCHCHK(c, launch_warpmap(c->d_warp_occ, c->d_warp_part, *p, c->d_warpmap, c->stream));
    c->have_warpmap = true;
    return VCT_OK;
}

int vct_normalize_inject(vct_ctx *c, const vct_inject_params *p) {
    if (!c || !p) return fail(c, VCT_ERR_INVALID, "vct_normalize_inject: NULL argument");
    if (!(p->voxel_set_opacity >= 0.0f && p->voxel_set_opacity <= 1.0f))
        return fail(c, VCT_ERR_INVALID, "voxel_set_opacity %f out of [0, 1]", (double)p->voxel_set_opacity);
    if (p->temporal_filter && !(p->temporal_decay >= 0.0f && p->temporal_decay <= 1.0f))
        return fail(c, VCT_ERR_INVALID, "temporal_decay %f out of [0, 1]", (double)p->temporal_decay);
    if (!c->have_voxels) return fail(c, VCT_ERR_STATE, "vct_normalize_inject: voxelize first");
    if (p->warp_texture && !c->have_warpmap)
        return fail(c, VCT_ERR_STATE, "vct_normalize_inject: warp_texture without vct_warpmap");
    PassTimer pt(c, P_RADIANCE);
    Extent ext;
    memcpy(ext.center, c->desc.center, 12);
    memcpy(ext.vmin, c->desc.vmin, 12);
    memcpy(ext.vmax, c->desc.vmax, 12);
    LAUNCHCHK(c, launch_inject_scatter(c->hlights.data(), c->d_lights, (int)c->lights.size(), p->light_mask,
                                       c->desc.dim, ext, p->warp_texture ? c->d_warpmap : nullptr, c->d_litmask,
                                       c->stream));
    c->albedo_zero_valid = false;
    LAUNCHCHK(c, launch_normalize_inject(c->d_grid_albedo, c->d_litmask, c->desc.dim, p->voxel_set_opacity,
                                         p->temporal_filter != 0, p->temporal_decay, c->pyr, c->d_stats,
                                         c->d_partials, c->stream));
    c->albedo_zero_valid = true;
    // only the first injection after the voxelizer sees its words (a later one
    // reads them after the opacity rewrite)
    c->grid_zero_valid = c->grids_std_vox;
    c->grids_std_vox = false;
    c->have_radiance = true;
    c->have_mips = false;
    // a plain injection rewrites every texel of level 0 from unorm8 albedo (no
    // negative values); the temporal filter keeps decay * the previous value of
    // every unlit voxel, so a signed level 0 from vct_upload stays signed
    if (!p->temporal_filter) c->pyr_nonneg = true;
    return VCT_OK;
}

int vct_normalize_inject_noalloc(vct_ctx *c, const vct_inject_params *p) {
    if (!c || !p) return fail(c, VCT_ERR_INVALID, "vct_normalize_inject_noalloc: NULL argument");
    if (c->desc.dim > kNoAllocMaxDim)
        return fail(c, VCT_ERR_INVALID,
                    "vct_normalize_inject_noalloc: voxel dim %d exceeds the static scratch (dim <= %d)",
                    c->desc.dim, kNoAllocMaxDim);
    if (!(p->voxel_set_opacity >= 0.0f && p->voxel_set_opacity <= 1.0f))
        return fail(c, VCT_ERR_INVALID, "voxel_set_opacity %f out of [0, 1]", (double)p->voxel_set_opacity);
    if (p->temporal_filter && !(p->temporal_decay >= 0.0f && p->temporal_decay <= 1.0f))
        return fail(c, VCT_ERR_INVALID, "temporal_decay %f out of [0, 1]", (double)p->temporal_decay);
    if (!c->have_voxels) return fail(c, VCT_ERR_STATE, "vct_normalize_inject_noalloc: voxelize first");
    if (p->warp_texture && !c->have_warpmap)
        return fail(c, VCT_ERR_STATE, "vct_normalize_inject_noalloc: warp_texture without vct_warpmap");
    const NoAllocScratch sc = noalloc_scratch(c->desc.device);
    if (!sc.litmask) return fail(c, VCT_ERR_STATE, "vct_normalize_inject_noalloc: scratch not set up");
    PassTimer pt(c, P_RADIANCE);
    Extent ext;
    memcpy(ext.center, c->desc.center, 12);
    memcpy(ext.vmin, c->desc.vmin, 12);
    memcpy(ext.vmax, c->desc.vmax, 12);
    // the same kernels as vct_normalize_inject, with the lit mask (cleared by a
    // fill kernel inside the scatter launch) and the partials in the scratch
    LAUNCHCHK(c, launch_inject_scatter(c->hlights.data(), c->d_lights, (int)c->lights.size(), p->light_mask,
                                       c->desc.dim, ext, p->warp_texture ? c->d_warpmap : nullptr, sc.litmask,
                                       c->stream));
    c->albedo_zero_valid = false;
    LAUNCHCHK(c, launch_normalize_inject(c->d_grid_albedo, sc.litmask, c->desc.dim, p->voxel_set_opacity,
                                         p->temporal_filter != 0, p->temporal_decay, c->pyr, c->d_stats,
                                         sc.partials, c->stream));
    c->albedo_zero_valid = true;
    // only the first injection after the voxelizer sees its words (a later one
    // reads them after the opacity rewrite)
    c->grid_zero_valid = c->grids_std_vox;
    c->grids_std_vox = false;
    c->have_radiance = true;
    c->have_mips = false;
    if (!p->temporal_filter) c->pyr_nonneg = true;
    return VCT_OK;
}

int vct_fill_holes(vct_ctx *c) {
    if (!c) return fail(nullptr, VCT_ERR_INVALID, "ctx is NULL");
    if (!c->have_radiance) return fail(c, VCT_ERR_STATE, "vct_fill_holes: inject first");
    const LevelInfo &L0 = c->pyr.lv[0];
    if (!c->d_fill_tmp) {  // level 0's layout; its fp16-pair border halves stay 0
        const size_t bytes = (size_t)L0.nelem * 16;
        HIPCHK(c, dalloc((char **)&c->d_fill_tmp, bytes));
        HIPCHK(c, hipMemsetAsync(c->d_fill_tmp, 0, bytes, c->stream));
    }
    PassTimer pt(c, P_FILL);
    LAUNCHCHK(c, launch_fill_holes(c->pyr, c->d_fill_tmp, c->stream));
    c->have_mips = false;
    return VCT_OK;
}

int vct_mip(vct_ctx *c, const vct_mip_params *p) {
    if (!c || !p) return fail(c, VCT_ERR_INVALID, "vct_mip: NULL argument");
    if (p->kernel_mode != 0)
        return fail(c, VCT_ERR_INVALID, "vct_mip: kernel_mode %d unsupported (only KERNEL_BOX2 = 0 is live)",
                    p->kernel_mode);
    if (!c->have_radiance) return fail(c, VCT_ERR_STATE, "vct_mip: inject first");
    PassTimer pt(c, P_MIP);
    // the voxelColor chain, filtered by the same mipmap pass (Application.cpp:
    // 904-917), depends only on the albedo grid: it runs on the aux stream
    // beside the radiance chain, from a fork after everything enqueued so far
    // to a join before anything enqueued later
    if (p->albedo_chain) {
        HIPCHK(c, hipEventRecord(c->fork, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->aux, c->fork, 0));
        LAUNCHCHK(c, launch_albedo_chain(c->d_grid_albedo, c->albedo_zero_valid ? c->pyr.zero0 : nullptr, c->albedo_lv,
                                         c->desc.dim, c->desc.levels, c->aux));
        HIPCHK(c, hipEventRecord(c->join, c->aux));
    }
    LAUNCHCHK(c, launch_mip(c->pyr, c->stream));
    if (p->albedo_chain) HIPCHK(c, hipStreamWaitEvent(c->stream, c->join, 0));
    c->have_mips = true;
    return VCT_OK;
}

static int ensure_camera(vct_ctx *c, int w, int h) {
    if (w == c->gw && h == c->gh && c->d_gbuf) return VCT_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c->d_gbuf);
    dfree(c->d_vis);
    dfree(c->d_ind);
    dfree(c->d_spec);
    dfree(c->d_color);
    dfree(c->d_reference);  // sized for the old image: the next vct_reference_gather allocates it anew
    dfree(c->d_tile_count);  // likewise for the next vct_trace_learned
    const size_t np = (size_t)w * h;
    HIPCHK(c, dalloc(&c->d_gbuf, np * sizeof(vct_gbuffer_texel)));
    HIPCHK(c, dalloc(&c->d_vis, np * 8));
    HIPCHK(c, dalloc(&c->d_ind, np * 16));
    HIPCHK(c, dalloc(&c->d_spec, np * 16));
    HIPCHK(c, dalloc(&c->d_color, np * 16));
    HIPCHK(c, hipMemset(c->d_color, 0, np * 16));
    HIPCHK(c, hipMemset(c->d_gbuf, 0, np * sizeof(vct_gbuffer_texel)));
    HIPCHK(c, hipMemset(c->d_ind, 0, np * 16));
    HIPCHK(c, hipMemset(c->d_spec, 0, np * 16));
    c->gw = w;
    c->gh = h;
    c->have_gbuf = false;
    return VCT_OK;
}

int vct_gbuffer(vct_ctx *c, const vct_camera *cam) {
    if (!c || !cam) return fail(c, VCT_ERR_INVALID, "vct_gbuffer: NULL argument");
    if (cam->width < 1 || cam->height < 1 || cam->width > 8192 || cam->height > 8192)
        return fail(c, VCT_ERR_INVALID, "camera size %dx%d out of [1, 8192]", cam->width, cam->height);
    if (!c->d_model) return fail(c, VCT_ERR_STATE, "vct_gbuffer: no scene");
    int rc = ensure_camera(c, cam->width, cam->height);
    if (rc) return rc;
    PassTimer pt(c, P_GBUFFER);
    LAUNCHCHK(c, launch_gbuffer(dev_scene(c), *cam, c->d_vis, c->d_gbuf, c->rs, c->stream));
    c->have_gbuf = true;
    return VCT_OK;
}

static TraceLaunch make_trace_launch(vct_ctx *c, const vct_trace_params *p);

int vct_trace(vct_ctx *c, const vct_trace_params *p) {
    if (!c || !p) return fail(c, VCT_ERR_INVALID, "vct_trace: NULL argument");
    if (!c->have_gbuf) return fail(c, VCT_ERR_STATE, "vct_trace: no G-buffer (call vct_gbuffer or upload one)");
    if (!c->have_mips) return fail(c, VCT_ERR_STATE, "vct_trace: mip chain not built");
    if (p->diffuse.steps < 0 || p->diffuse.steps > 4096 || p->specular.steps < 0 || p->specular.steps > 4096)
        return fail(c, VCT_ERR_INVALID, "cone steps out of [0, 4096]");
    if (p->warp_texture && !c->have_warpmap) return fail(c, VCT_ERR_STATE, "vct_trace: warp_texture without vct_warpmap");
    const TraceLaunch t = make_trace_launch(c, p);
    PassTimer pt(c, P_TRACE);
    LAUNCHCHK(c, launch_trace(c->pyr, t, c->stream));
    c->have_trace = true;
    c->trace_occluded = p->draw_occlusion != 0;
    c->trace_spec = p->enable_reflections != 0;
    return VCT_OK;
}

int vct_compose(vct_ctx *c, const vct_compose_params *p) {
    if (!c || !p) return fail(c, VCT_ERR_INVALID, "vct_compose: NULL argument");
    if (!c->have_gbuf) return fail(c, VCT_ERR_STATE, "vct_compose: no G-buffer");
    if (p->enable_indirect && !c->have_trace)
        return fail(c, VCT_ERR_STATE, "vct_compose: enable_indirect needs a vct_trace of the current G-buffer");
    if (p->enable_indirect && p->enable_reflections && !c->trace_spec)
        return fail(c, VCT_ERR_STATE, "vct_compose: enable_reflections needs a vct_trace with reflections");
    PassTimer pt(c, P_COMPOSE);
    LAUNCHCHK(c, launch_compose(c->d_gbuf, c->d_ind, c->d_spec, c->d_color, c->gw * c->gh,
                                c->lights.empty() ? nullptr : c->d_lights, (int)c->lights.size(), c->has_rough ? 1 : 0,
                                c->trace_occluded ? 1 : 0, *p, c->stream));
    return VCT_OK;
}

int vct_sync(vct_ctx *c) {
    if (!c) return fail(nullptr, VCT_ERR_INVALID, "ctx is NULL");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VCT_OK;
}

int vct_get_stats(vct_ctx *c, vct_stats *out) {
    if (!c || !out) return fail(c, VCT_ERR_INVALID, "vct_get_stats: NULL argument");
    HIPCHK(c, hipMemcpyAsync(out, c->d_stats, sizeof(vct_stats), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VCT_OK;
}

int vct_set_timing(vct_ctx *c, int enable) {
    if (!c) return fail(nullptr, VCT_ERR_INVALID, "ctx is NULL");
    c->timing = enable != 0;
    return VCT_OK;
}

int vct_get_pass_times(vct_ctx *c, vct_pass_times *out) {
    if (!c || !out) return fail(c, VCT_ERR_INVALID, "vct_get_pass_times: NULL argument");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double ms[P_COUNT] = {0};
    for (auto &tp : c->pending) {
        float v = 0.0f;
        if (hipEventElapsedTime(&v, tp.a, tp.b) == hipSuccess) ms[tp.pass] += v;
        c->ev_pool.push_back(tp.a);
        c->ev_pool.push_back(tp.b);
    }
    c->pending.clear();
    out->shadowmap_ms = ms[P_SHADOW];
    out->voxelize_ms = ms[P_VOXELIZE];
    out->radiance_ms = ms[P_RADIANCE];
    out->fill_ms = ms[P_FILL];
    out->mipmap_ms = ms[P_MIP];
    out->gbuffer_ms = ms[P_GBUFFER];
    out->render_ms = ms[P_TRACE];
    out->learned_ms = ms[P_LEARNED];
    out->compose_ms = ms[P_COMPOSE];
    out->warpmap_ms = ms[P_WARP];
    return VCT_OK;
}

static int decode_level(const vct_ctx *c, int level, int *l, int *face) {
    *l = level & 15;
    *face = level >> 4;
    if (*l < 0 || *l >= c->pyr.nlevels) return 0;
    if (*face < 0 || *face >= c->pyr.lv[*l].nfaces) return 0;
    return 1;
}

size_t vct_buffer_bytes(const vct_ctx *c, int what, int level) {
    if (!c) return 0;
    const size_t dim = (size_t)c->desc.dim, nv = dim * dim * dim;
    int l, f;
    switch (what) {
    case VCT_BUF_ALBEDO:
        if (level == 0) return nv * 4;
        if (level < 0 || level >= c->desc.levels) return 0;
        {
            const size_t sz = (size_t)c->pyr.lv[level].size;
            return sz * sz * sz * 4;
        }
    case VCT_BUF_NORMAL: return level == 0 ? nv * 4 : 0;
    case VCT_BUF_LITMASK: return nv;
    case VCT_BUF_RADIANCE:
        if (!decode_level(c, level, &l, &f)) return 0;
        return (size_t)c->pyr.lv[l].size * c->pyr.lv[l].size * c->pyr.lv[l].size * 16;
    case VCT_BUF_RADIANCE_RAW: {
        if (!decode_level(c, level, &l, &f)) return 0;
        const LevelInfo &L = c->pyr.lv[l];
        return (size_t)L.nelem * (L.fmt == 2 ? 4 : 16);
    }
    case VCT_BUF_SHADOWMAP:
        if (level < 0 || level >= (int)c->lights.size() || !c->lights[level].shadow_caster) return 0;
        return (size_t)c->lights[level].shadow_size * c->lights[level].shadow_size * 4;
    case VCT_BUF_GBUFFER: return (size_t)c->gw * c->gh * sizeof(vct_gbuffer_texel);
    case VCT_BUF_INDIRECT:
    case VCT_BUF_SPECULAR:
    case VCT_BUF_COLOR:
    case VCT_BUF_REFERENCE: return (size_t)c->gw * c->gh * 16;
    case VCT_BUF_WARPMAP: return c->have_warpmap ? (size_t)kWarpDim * kWarpDim * kWarpDim * 16 : 0;
    case VCT_BUF_WARP_OCCUPANCY: return c->have_warpmap ? (size_t)kWarpDim * kWarpDim * kWarpDim * 4 : 0;
    case VCT_BUF_LABELS: return (size_t)c->npoints * 16;
    case VCT_BUF_FEATURES: return (size_t)c->npoints * kMlpIn * 2;
    case VCT_BUF_WEIGHTS:
    case VCT_BUF_GRADIENT: return (size_t)(kMlpParams + 1) * 4;
    default: return 0;
    }
}

static int ensure_stage(vct_ctx *c, size_t bytes) {
    if (bytes <= c->stage_bytes) return VCT_OK;
    dfree(c->d_stage);
    HIPCHK(c, dalloc(&c->d_stage, bytes));
    c->stage_bytes = bytes;
    return VCT_OK;
}

int vct_download(vct_ctx *c, int what, int level, void *host, size_t bytes) {
    if (!c || !host) return fail(c, VCT_ERR_INVALID, "vct_download: NULL argument");
    const size_t need = vct_buffer_bytes(c, what, level);
    if (need == 0) return fail(c, VCT_ERR_INVALID, "vct_download: buffer %d level %d does not exist", what, level);
    if (bytes != need) return fail(c, VCT_ERR_INVALID, "vct_download: %zu bytes given, %zu needed", bytes, need);
    const void *src = nullptr;
    int l, f;
    switch (what) {
    case VCT_BUF_ALBEDO:
    case VCT_BUF_NORMAL: {  // bricked on the device, linear x-fastest for the caller
        if (what == VCT_BUF_ALBEDO && level > 0) {  // the albedo chain is linear
            src = c->albedo_lv[level];
            break;
        }
        int rc = ensure_stage(c, need);
        if (rc) return rc;
        LAUNCHCHK(c, launch_grid_convert(what == VCT_BUF_ALBEDO ? c->d_grid_albedo : c->d_grid_normal,
                                         (uint32_t *)c->d_stage, c->desc.dim, true, c->stream));
        src = c->d_stage;
        break;
    }
    case VCT_BUF_LITMASK: src = c->d_litmask; break;
    case VCT_BUF_SHADOWMAP: src = c->d_shadow[level]; break;
    case VCT_BUF_GBUFFER: src = c->d_gbuf; break;
    case VCT_BUF_INDIRECT: src = c->d_ind; break;
    case VCT_BUF_SPECULAR: src = c->d_spec; break;
    case VCT_BUF_COLOR: src = c->d_color; break;
    case VCT_BUF_WARPMAP: src = c->d_warpmap; break;
    case VCT_BUF_WARP_OCCUPANCY: src = c->d_warp_occ; break;
    case VCT_BUF_LABELS: src = c->d_labels; break;
    case VCT_BUF_FEATURES: src = c->d_X; break;
    case VCT_BUF_REFERENCE:
        if (!c->d_reference) return fail(c, VCT_ERR_STATE, "vct_download: no reference gather computed");
        src = c->d_reference;
        break;
    case VCT_BUF_GRADIENT: {
        if (!c->mlp.partials || c->last_train_blocks <= 0)
            return fail(c, VCT_ERR_STATE, "vct_download: no training step has run");
        std::vector<float> part((size_t)c->last_train_blocks * kMlpPartialStride);
        HIPCHK(c, hipMemcpyAsync(part.data(), c->mlp.partials, part.size() * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        float *o = (float *)host;
        for (int p = 0; p <= kMlpParams; ++p) {
            double s = 0.0;
            for (int b = 0; b < c->last_train_blocks; ++b) s += part[(size_t)b * kMlpPartialStride + p];
            o[p] = (float)s;
        }
        return VCT_OK;
    }
    case VCT_BUF_WEIGHTS: {
        if (!c->mlp.master) return fail(c, VCT_ERR_STATE, "vct_download: no learned weights");
        HIPCHK(c, hipMemcpyAsync(host, c->mlp.master, (size_t)kMlpParams * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        ((float *)host)[kMlpParams] = c->mlp.out_scale;
        return VCT_OK;
    }
    case VCT_BUF_RADIANCE: {
        decode_level(c, level, &l, &f);
        int rc = ensure_stage(c, need);
        if (rc) return rc;
        LAUNCHCHK(c, launch_pyr_to_linear(c->pyr, l, f, c->d_stage, c->stream));
        src = c->d_stage;
        break;
    }
    case VCT_BUF_RADIANCE_RAW:
        decode_level(c, level, &l, &f);
        {  // the stored level, border bricks included
            const LevelInfo &L = c->pyr.lv[l];
            src = c->pyr.pyr + L.base - (size_t)L.pad * (L.fmt == 2 ? 4 : 16);
        }
        break;
    default: return fail(c, VCT_ERR_INVALID, "vct_download: unknown buffer %d", what);
    }
    HIPCHK(c, hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VCT_OK;
}

int vct_upload(vct_ctx *c, int what, int level, const void *host, size_t bytes) {
    if (!c || !host) return fail(c, VCT_ERR_INVALID, "vct_upload: NULL argument");
    const size_t need = vct_buffer_bytes(c, what, level);
    if (need == 0) return fail(c, VCT_ERR_INVALID, "vct_upload: buffer %d level %d does not exist", what, level);
    if (bytes != need) return fail(c, VCT_ERR_INVALID, "vct_upload: %zu bytes given, %zu needed", bytes, need);
    void *dst = nullptr;
    switch (what) {
    case VCT_BUF_WEIGHTS: {
        int rc = ensure_mlp(c);
        if (rc) return rc;
        const float *w = (const float *)host;
        if (!(w[kMlpParams] > 0.0f)) return fail(c, VCT_ERR_INVALID, "vct_upload: output scale must be > 0");
        HIPCHK(c, hipMemcpyAsync(c->mlp.master, w, (size_t)kMlpParams * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemsetAsync(c->mlp.m1, 0, (size_t)kMlpParams * 4, c->stream));
        HIPCHK(c, hipMemsetAsync(c->mlp.m2, 0, (size_t)kMlpParams * 4, c->stream));
        c->adam_t = 0;
        LAUNCHCHK(c, launch_emit_operands(c->mlp, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->mlp.out_scale = w[kMlpParams];
        c->have_weights = true;
        return VCT_OK;
    }
    case VCT_BUF_ALBEDO:
    case VCT_BUF_NORMAL: {
        if (level != 0) return fail(c, VCT_ERR_INVALID, "vct_upload: only level 0 of the albedo grid can be uploaded");
        int rc = ensure_stage(c, need);
        if (rc) return rc;
        HIPCHK(c, hipMemcpyAsync(c->d_stage, host, bytes, hipMemcpyHostToDevice, c->stream));
        LAUNCHCHK(c, launch_grid_convert((const uint32_t *)c->d_stage,
                                         what == VCT_BUF_ALBEDO ? c->d_grid_albedo : c->d_grid_normal, c->desc.dim,
                                         false, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->grids_std_vox = c->grid_zero_valid = false;
        if (what == VCT_BUF_ALBEDO) {
            c->have_voxels = true;
            c->albedo_zero_valid = false;
        }
        return VCT_OK;
    }
    case VCT_BUF_LITMASK: dst = c->d_litmask; break;
    case VCT_BUF_SHADOWMAP: {
        HIPCHK(c, hipMemcpyAsync(c->d_shadow[level], host, bytes, hipMemcpyHostToDevice, c->stream));
        LAUNCHCHK(c, launch_shadow_quads(c->d_shadow[level], (int)c->lights[level].shadow_size, c->d_shadowq[level],
                                         c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return VCT_OK;
    }
    case VCT_BUF_GBUFFER: dst = c->d_gbuf; c->have_gbuf = true; break;
    case VCT_BUF_RADIANCE: {
        if (level != 0) return fail(c, VCT_ERR_INVALID, "vct_upload: only radiance level 0 can be uploaded");
        int rc = ensure_stage(c, need);
        if (rc) return rc;
        HIPCHK(c, hipMemcpyAsync(c->d_stage, host, bytes, hipMemcpyHostToDevice, c->stream));
        LAUNCHCHK(c, launch_linear_to_pyr0(c->pyr, c->d_stage, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->have_radiance = true;
        c->have_mips = false;
        c->pyr_nonneg = false;  // the caller's values may be negative
        return VCT_OK;
    }
    default: return fail(c, VCT_ERR_INVALID, "vct_upload: buffer %d cannot be uploaded", what);
    }
    HIPCHK(c, hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VCT_OK;
}


// ---------------------------------------------------------------- learned gather
static int ensure_mlp(vct_ctx *c) {
    if (c->mlp.master) return VCT_OK;
    MlpBuffers &B = c->mlp;
    HIPCHK(c, dalloc(&B.master, (size_t)kMlpParams * 4));
    HIPCHK(c, dalloc(&B.m1, (size_t)kMlpParams * 4));
    HIPCHK(c, dalloc(&B.m2, (size_t)kMlpParams * 4));
    HIPCHK(c, dalloc(&B.w1, (size_t)64 * kMlpIn * 2));
    HIPCHK(c, dalloc(&B.w2, 64 * 64 * 2));
    HIPCHK(c, dalloc(&B.w3, 64 * 64 * 2));
    HIPCHK(c, dalloc(&B.w4, 16 * 64 * 2));
    HIPCHK(c, dalloc(&B.w2t, 64 * 64 * 2));
    HIPCHK(c, dalloc(&B.w3t, 64 * 64 * 2));
    HIPCHK(c, dalloc(&B.w4t, 64 * 32 * 2));
    HIPCHK(c, hipMemset(B.w1, 0, (size_t)64 * kMlpIn * 2));
    HIPCHK(c, hipMemset(B.w4, 0, 16 * 64 * 2));
    HIPCHK(c, hipMemset(B.w4t, 0, 64 * 32 * 2));
    HIPCHK(c, hipMemset(B.master, 0, (size_t)kMlpParams * 4));
    return VCT_OK;
}

int vct_learned_set_points(vct_ctx *c, const vct_gbuffer_texel *pts, int n) {
    if (!c || !pts) return fail(c, VCT_ERR_INVALID, "vct_learned_set_points: NULL argument");
    if (n < 64 || n > (1 << 24)) return fail(c, VCT_ERR_INVALID, "vct_learned_set_points: %d points (64..2^24)", n);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c->d_points);
    dfree(c->d_labels);
    dfree(c->d_X);
    HIPCHK(c, dalloc(&c->d_points, (size_t)n * sizeof(vct_gbuffer_texel)));
    HIPCHK(c, dalloc(&c->d_labels, (size_t)n * 16));
    HIPCHK(c, dalloc(&c->d_X, (size_t)n * kMlpIn * 2));
    HIPCHK(c, hipMemcpy(c->d_points, pts, (size_t)n * sizeof(vct_gbuffer_texel), hipMemcpyHostToDevice));
    c->npoints = n;
    c->have_prepared = false;
    return VCT_OK;
}

static Extent ctx_extent(const vct_ctx *c) {
    Extent e;
    memcpy(e.center, c->desc.center, 12);
    memcpy(e.vmin, c->desc.vmin, 12);
    memcpy(e.vmax, c->desc.vmax, 12);
    return e;
}

static int learned_ok(vct_ctx *c) {
    if (c->pyr.aniso) return fail(c, VCT_ERR_INVALID, "the learned gather needs an isotropic mip chain");
    if (c->pyr.nlevels < 3) return fail(c, VCT_ERR_INVALID, "the learned gather needs mip level 2");
    if (!c->have_mips) return fail(c, VCT_ERR_STATE, "the mip chain is not built");
    return VCT_OK;
}

int vct_learned_prepare(vct_ctx *c) {
    if (!c) return fail(nullptr, VCT_ERR_INVALID, "ctx is NULL");
    int rc = learned_ok(c);
    if (rc) return rc;
    if (!c->d_points) return fail(c, VCT_ERR_STATE, "vct_learned_prepare: no points");
    const Extent e = ctx_extent(c);
    LAUNCHCHK(c, launch_label64(c->pyr, c->d_points, c->npoints, e, c->d_labels, c->stream));
    LAUNCHCHK(c, launch_features(c->pyr, c->d_points, c->npoints, e, c->d_X, c->stream));
    // output scale: the mean label magnitude, so the network fits O(1) targets
    std::vector<float> lab((size_t)c->npoints * 4);
    HIPCHK(c, hipMemcpyAsync(lab.data(), c->d_labels, lab.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (int i = 0; i < c->npoints; ++i) s += fabs(lab[i * 4]) + fabs(lab[i * 4 + 1]) + fabs(lab[i * 4 + 2]);
    const double mean = s / (3.0 * c->npoints);
    c->mlp.out_scale = (float)(mean > 1e-6 ? mean : 1e-6);
    c->have_prepared = true;
    return VCT_OK;
}

int vct_learned_train(vct_ctx *c, const vct_learned_params *p, float *loss) {
    if (!c || !p) return fail(c, VCT_ERR_INVALID, "vct_learned_train: NULL argument");
    if (p->batch < 64 || p->batch % 64 || p->batch > 65536)
        return fail(c, VCT_ERR_INVALID, "batch %d must be a multiple of 64 in [64, 65536]", p->batch);
    if (p->steps < 0 || p->steps > 1000000) return fail(c, VCT_ERR_INVALID, "steps %d out of range", p->steps);
    if (!(p->lr > 0.0f)) return fail(c, VCT_ERR_INVALID, "learning rate must be > 0");
    if (!(p->lr_final >= 0.0f)) return fail(c, VCT_ERR_INVALID, "final learning rate must be >= 0");
    if (!c->have_prepared) return fail(c, VCT_ERR_STATE, "vct_learned_train: call vct_learned_prepare first");
    int rc = ensure_mlp(c);
    if (rc) return rc;
    MlpBuffers &B = c->mlp;
    const int nblk = p->batch / 64;
    if (nblk > B.max_blocks) {
        dfree(B.partials);
        HIPCHK(c, dalloc(&B.partials, (size_t)nblk * kMlpPartialStride * 4));
        B.max_blocks = nblk;
    }
    if (p->steps > c->loss_cap) {
        dfree(c->d_loss);
        HIPCHK(c, dalloc(&c->d_loss, (size_t)p->steps * 4));
        c->loss_cap = p->steps;
    }
    if (p->reset || !c->have_weights) {
        // He-uniform initialisation U(-sqrt(6/fan_in), sqrt(6/fan_in)), zero biases
        std::vector<float> w(kMlpParams, 0.0f);
        std::mt19937_64 rng(p->seed);
        const int fin[4] = {129, 64, 64, 64}, fout[4] = {64, 64, 64, 3};
        size_t o = 0;
        for (int l = 0; l < 4; ++l) {
            std::uniform_real_distribution<double> U(-sqrt(6.0 / fin[l]), sqrt(6.0 / fin[l]));
            for (int i = 0; i < fin[l] * fout[l]; ++i) w[o++] = (float)U(rng);
            o += fout[l];
        }
        HIPCHK(c, hipMemcpy(B.master, w.data(), (size_t)kMlpParams * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemset(B.m1, 0, (size_t)kMlpParams * 4));
        HIPCHK(c, hipMemset(B.m2, 0, (size_t)kMlpParams * 4));
        LAUNCHCHK(c, launch_emit_operands(B, c->stream));
        c->have_weights = true;
        c->adam_t = 0;
    }
    PassTimer pt(c, P_LEARNED);
    if (p->steps > 0) c->last_train_blocks = nblk;
    const bool cosine = p->lr_final > 0.0f && p->lr_final != p->lr && p->steps > 1;
    for (int s = 0; s < p->steps; ++s) {
        const float lr = cosine ? p->lr_final + 0.5f * (p->lr - p->lr_final) *
                                                     (1.0f + cosf(3.14159265f * (float)s / (float)(p->steps - 1)))
                                : p->lr;
        LAUNCHCHK(c, launch_mlp_train_step(B, c->d_X, c->d_labels, c->npoints, p->batch, p->seed, ++c->adam_t, lr,
                                           p->beta1, p->beta2, p->eps, c->d_loss + s, c->stream));
    }
    pt.stop();
    if (loss && p->steps) HIPCHK(c, hipMemcpyAsync(loss, c->d_loss, (size_t)p->steps * 4, hipMemcpyDeviceToHost,
                                                   c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VCT_OK;
}

static TraceLaunch make_trace_launch(vct_ctx *c, const vct_trace_params *p) {
    TraceLaunch t;
    t.gbuf = c->d_gbuf;
    t.width = c->gw;
    t.height = c->gh;
    t.out_ind = c->d_ind;
    t.out_spec = c->d_spec;
    t.enable_reflections = p->enable_reflections;
    t.angle_from_roughness = p->cone_angle_from_roughness;
    t.draw_occlusion = p->draw_occlusion;
    t.has_roughness = c->has_rough;
    t.ambient_scale = p->ambient_scale;
    t.reflect_scale = p->reflect_scale;
    memcpy(t.eye, p->eye, 12);
    t.d_steps = p->diffuse.steps;
    t.d_tan_half = tanf(p->diffuse.cone_angle / 2.0f);
    t.d_bias = p->diffuse.bias;
    t.d_h0 = p->diffuse.cone_initial_height;
    t.d_lod_offset = p->diffuse.lod_offset;
    t.s_steps = p->specular.steps;
    t.s_angle = p->specular.cone_angle;
    t.s_bias = p->specular.bias;
    t.s_h0 = p->specular.cone_initial_height;
    t.s_lod_offset = p->specular.lod_offset;
    t.ext = ctx_extent(c);
    t.warp = p->warp_texture ? c->d_warpmap : nullptr;
    t.nested = c->pyr_nonneg ? 1 : 0;
    // the diffuse schedule, step by step as traceCone computes it
    // (phong.frag:145-168): r = h tan(theta/2), lambda = log2(max(1, 2r)) +
    // lodOffset, the level class of lambda (LINEAR_MIPMAP_LINEAR with the clamp
    // to [0, q]), h += r
    memset(&t.d_sched, 0, sizeof(t.d_sched));
    if (t.d_steps > 0 && t.d_steps <= kMaxSchedule) {
        const int q = c->desc.levels - 1;
        float h = t.d_h0;
        for (int i = 0; i < t.d_steps; ++i) {
            const float r = h * t.d_tan_half;
            const float m = 2.0f * r;
            const float lambda = log2f(m > 1.0f ? m : 1.0f) + t.d_lod_offset;
            int cls;
            float fr = 0.0f;
            if (lambda <= 0.0f) cls = -1;
            else if (!(lambda < (float)q)) cls = q;
            else {
                const float fl = floorf(lambda);
                cls = std::min(std::max((int)fl, 0), q - 1);
                fr = lambda - fl;
            }
            t.d_sched.h[i] = h;
            t.d_sched.cls[i] = cls;
            t.d_sched.fr[i] = fr;
            h = h + r;
        }
        t.d_sched.n = t.d_steps;
    }
    return t;
}

int vct_trace_learned(vct_ctx *c, const vct_trace_params *p, float thr, int *fallback) {
    if (!c || !p) return fail(c, VCT_ERR_INVALID, "vct_trace_learned: NULL argument");
    int rc = learned_ok(c);
    if (rc) return rc;
    if (!c->have_gbuf) return fail(c, VCT_ERR_STATE, "vct_trace_learned: no G-buffer");
    if (!c->have_weights) return fail(c, VCT_ERR_STATE, "vct_trace_learned: no trained weights");
    if (p->warp_texture)
        return fail(c, VCT_ERR_INVALID, "vct_trace_learned: the learned gather is trained on the unwarped grid");
    if (!c->d_tile_count)  // sized for the image: ensure_camera frees it
        HIPCHK(c, dalloc(&c->d_tile_count, 4 * (1 + (size_t)((c->gw + 7) / 8) * ((c->gh + 7) / 8))));
    TraceLaunch t = make_trace_launch(c, p);
    t.enable_reflections = 0;
    PassTimer pt(c, P_LEARNED);
    LAUNCHCHK(c, launch_learned(c->mlp, c->pyr, t, thr, c->d_tile_count, c->stream));
    pt.stop();
    if (fallback && !(thr < __builtin_huge_valf())) {
        *fallback = 0;  // no tile can exceed an infinite threshold: nothing was listed
    } else if (fallback) {
        uint32_t n = 0;
        HIPCHK(c, hipMemcpyAsync(&n, c->d_tile_count, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *fallback = (int)n;
    }
    return VCT_OK;
}

int vct_reference_gather(vct_ctx *c) {
    if (!c) return fail(nullptr, VCT_ERR_INVALID, "ctx is NULL");
    int rc = learned_ok(c);
    if (rc) return rc;
    if (!c->have_gbuf) return fail(c, VCT_ERR_STATE, "vct_reference_gather: no G-buffer");
    if (!c->d_reference) HIPCHK(c, dalloc(&c->d_reference, (size_t)c->gw * c->gh * 16));
    LAUNCHCHK(c, launch_label64(c->pyr, c->d_gbuf, c->gw * c->gh, ctx_extent(c), c->d_reference, c->stream));
    return VCT_OK;
}

}  // extern "C"
