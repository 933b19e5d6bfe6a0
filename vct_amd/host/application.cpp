// application.cpp -- headless application layer (include/vct_app.h).
//
// Builds the seeded synthetic scenes, owns the per-frame uniforms and runs the
// pass order of Application::render (src/Application.cpp:196-1085) through
// the C ABI of include/vct.h.  The scenes replace the reference's Sponza actor
// (Application.cpp:93-136): triangle soups with per-triangle RGBA8 albedo
// (the 1x1 diffuse texture a material would bind), face or smooth normals and
// per-triangle tangents.  The tangent is what the Mesh.cpp:174-197 rule gives
// for the uv layout the scene implies: box faces carry one uv square per face
// (u along the face's first in-plane axis), so both triangles of a face share
// the tangent of that axis; sphere triangles use the canonical in-plane axis
// normalize((0,1,0) x n) (normalize((1,0,0) x n) near the poles).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <cstring>
#include <string>
#include <vector>

#include "glmath.h"
#include "vct_app.h"

using namespace vcthost;

namespace {

// SplitMix64 stream; floats from the top 24 bits.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * uni(); }
    double normal() {
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Mesh {
    std::vector<float> pos, nrm, tan, rough;
    std::vector<float> vtan;  // per-corner (vertex) tangents, [ntri][3][3], or empty
    std::vector<float> uv;    // per-corner texture coordinates, [ntri][3][2], or empty
    std::vector<int32_t> texid;  // per-triangle diffuse map, or empty
    std::vector<uint32_t> albedo, obj;
    int nobj = 0;
    int ntri() const { return (int)albedo.size(); }
};

uint32_t random_albedo(Rng &r) {
    uint32_t v = 0;
    for (int k = 0; k < 3; ++k) {
        double u = r.uni(0.2, 0.9);
        uint32_t q = (uint32_t)std::lrint(u * 255.0);
        v |= (q & 0xffu) << (8 * k);
    }
    return v | 0xff000000u;
}

void push_tri(Mesh &m, const double p[3][3], const double n[3][3], const double tangent[3], uint32_t albedo,
              uint32_t obj) {
    for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) {
            m.pos.push_back((float)p[v][k]);
            m.nrm.push_back((float)n[v][k]);
        }
    const double l = std::sqrt(tangent[0] * tangent[0] + tangent[1] * tangent[1] + tangent[2] * tangent[2]);
    for (int k = 0; k < 3; ++k) m.tan.push_back((float)(tangent[k] / l));
    m.albedo.push_back(albedo);
    m.obj.push_back(obj);
}

// Axis-aligned box centred on the object origin; outward (or inward for the
// room) face normals, counter-clockwise as seen from the side the normal
// points to.
void add_box(Mesh &m, const double h[3], bool inward, uint32_t obj, Rng &r) {
    for (int a = 0; a < 3; ++a)
        for (int sgn = -1; sgn <= 1; sgn += 2) {
            const int u = (a + 1) % 3, v = (a + 2) % 3;
            double corner[4][3];
            const double cu[4] = {-1, 1, 1, -1}, cv[4] = {-1, -1, 1, 1};
            for (int i = 0; i < 4; ++i) {
                corner[i][a] = sgn * h[a];
                corner[i][u] = cu[i] * h[u];
                corner[i][v] = cv[i] * h[v];
            }
            bool flip = (sgn < 0) != inward;
            double nrm[3] = {0, 0, 0};
            nrm[a] = inward ? -sgn : sgn;
            double tng[3] = {0, 0, 0};  // uv (cu+1)/2, (cv+1)/2: dp/du along axis u
            tng[u] = 1.0;
            const int q[2][3] = {{0, 1, 2}, {0, 2, 3}};
            for (int t = 0; t < 2; ++t) {
                double p[3][3], n[3][3];
                for (int k = 0; k < 3; ++k) {
                    int idx = q[t][flip ? (k == 0 ? 0 : 3 - k) : k];
                    std::memcpy(p[k], corner[idx], sizeof(p[k]));
                    std::memcpy(n[k], nrm, sizeof(n[k]));
                }
                push_tri(m, p, n, tng, random_albedo(r), obj);
            }
        }
}

void add_icosphere(Mesh &m, double radius, int subdiv, uint32_t obj, Rng &r) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<std::array<double, 3>> V = {{-1, t, 0}, {1, t, 0},  {-1, -t, 0}, {1, -t, 0},
                                            {0, -1, t}, {0, 1, t},  {0, -1, -t}, {0, 1, -t},
                                            {t, 0, -1}, {t, 0, 1},  {-t, 0, -1}, {-t, 0, 1}};
    for (auto &v : V) {
        double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (auto &c : v) c /= l;
    }
    std::vector<std::array<int, 3>> F = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                         {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                         {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                         {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    for (int s = 0; s < subdiv; ++s) {
        std::vector<std::array<int, 3>> F2;
        F2.reserve(F.size() * 4);
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            std::array<double, 3> p;
            for (int k = 0; k < 3; ++k) p[k] = (V[a][k] + V[b][k]) * 0.5;
            double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            for (auto &c : p) c /= l;
            V.push_back(p);
            mid[key] = (int)V.size() - 1;
            return (int)V.size() - 1;
        };
        for (auto &f : F) {
            int ab = midpoint(f[0], f[1]), bc = midpoint(f[1], f[2]), ca = midpoint(f[2], f[0]);
            F2.push_back({f[0], ab, ca});
            F2.push_back({f[1], bc, ab});
            F2.push_back({f[2], ca, bc});
            F2.push_back({ab, bc, ca});
        }
        F.swap(F2);
    }
    for (auto &f : F) {
        double p[3][3], n[3][3];
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) {
                n[k][c] = V[f[k]][c];
                p[k][c] = V[f[k]][c] * radius;
            }
        // make the winding counter-clockwise seen from outside
        double e1[3], e2[3], cr[3];
        for (int c = 0; c < 3; ++c) {
            e1[c] = p[1][c] - p[0][c];
            e2[c] = p[2][c] - p[0][c];
        }
        cr[0] = e1[1] * e2[2] - e1[2] * e2[1];
        cr[1] = e1[2] * e2[0] - e1[0] * e2[2];
        cr[2] = e1[0] * e2[1] - e1[1] * e2[0];
        double cen = cr[0] * (p[0][0] + p[1][0] + p[2][0]) + cr[1] * (p[0][1] + p[1][1] + p[2][1]) +
                     cr[2] * (p[0][2] + p[1][2] + p[2][2]);
        if (cen < 0) {
            for (int c = 0; c < 3; ++c) {
                std::swap(p[1][c], p[2][c]);
                std::swap(n[1][c], n[2][c]);
            }
        }
        double fn[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) fn[c] += n[k][c];
        const double fl = std::sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
        for (auto &c : fn) c /= fl;
        // (0,1,0) x n, or (1,0,0) x n within ~25 degrees of the poles
        double tng[3];
        if (std::fabs(fn[1]) < 0.9) {
            tng[0] = fn[2];
            tng[1] = 0.0;
            tng[2] = -fn[0];
        } else {
            tng[0] = 0.0;
            tng[1] = -fn[2];
            tng[2] = fn[1];
        }
        push_tri(m, p, n, tng, random_albedo(r), obj);
    }
}

// Rigid transform T(c) * R(axis, angle) as a column-major float matrix.
void rigid(const double c[3], const double axis[3], double angle, float out[16]) {
    const double s = std::sin(angle), co = std::cos(angle), ic = 1.0 - co;
    const double x = axis[0], y = axis[1], z = axis[2];
    double R[3][3] = {{co + x * x * ic, x * y * ic - z * s, x * z * ic + y * s},
                      {y * x * ic + z * s, co + y * y * ic, y * z * ic - x * s},
                      {z * x * ic - y * s, z * y * ic + x * s, co + z * z * ic}};
    identity(out);
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) out[col * 4 + row] = (float)R[row][col];
    out[12] = (float)c[0];
    out[13] = (float)c[1];
    out[14] = (float)c[2];
}

double reflect_into(double p, double B) {
    double u = std::fmod(p + B, 4.0 * B);
    if (u < 0) u += 4.0 * B;
    return u < 2.0 * B ? u - B : 3.0 * B - u;
}

struct ObjMotion {
    double c0[3], vel[3], axis[3], omega;
};

}  // namespace

struct vct_app {
    vct_app_config cfg{};
    vct_ctx *ctx = nullptr;
    std::string err;
    Mesh mesh;
    std::vector<ObjMotion> motion;
    bool dynamic = false;
    std::vector<vct_light> lights;
    float extent_min[3], extent_max[3], center[3];
    vct_camera cam{};
    vct_trace_params trace{};
    std::vector<float> model;  // scratch
    std::vector<vct_texture> textures;  // the scene's diffuse maps
    std::vector<std::vector<uint32_t>> tex_texels;
    // where seeded cameras may stand (vct_app_camera): not inside these boxes
    std::vector<std::array<double, 6>> boxes;
    double cam_range = 0.0;
    float cam_znear = 0.05f, cam_zfar = 10.0f;
};

static int app_fail(vct_app *a, int code, const std::string &msg) {
    if (a) a->err = msg;
    return code;
}

static void make_dir_light(Rng &r, double bound, int sm, vct_light &L) {
    std::memset(&L, 0, sizeof(L));
    const double az = r.uni(0.0, 6.283185307179586), el = r.uni(0.5235987755982988, 1.3089969389957472);
    const float dir[3] = {(float)(-std::cos(el) * std::cos(az)), (float)(-std::sin(el)),
                          (float)(-std::cos(el) * std::sin(az))};
    const float dist = (float)(bound * 1.7);
    const float pos[3] = {-dir[0] * dist, -dir[1] * dist, -dir[2] * dist};
    std::memcpy(L.position, pos, 12);
    std::memcpy(L.direction, dir, 12);
    L.color[0] = L.color[1] = L.color[2] = 1.0f;
    L.intensity = 1.0f;
    L.range = 5.0f;
    L.type = VCT_LIGHT_DIRECTIONAL;
    L.enabled = 1;
    L.shadow_caster = 1;
    L.shadow_size = (uint32_t)sm;
    L.perspective = 0;
    // lp = ortho(-b, b, -b, b, 0, far); lv = lookAt(pos, pos + dir, up)  (Application.cpp:207-210)
    float lp[16], lv[16];
    const float b = (float)bound;
    ortho(-b, b, -b, b, 0.0f, 2.0f * dist, lp);
    const float tgt[3] = {pos[0] + dir[0], pos[1] + dir[1], pos[2] + dir[2]};
    const float up[3] = {0.0f, 1.0f, 0.0f};
    lookat(pos, tgt, up, lv);
    mul4(lp, lv, L.ls);
    inverse4(L.ls, L.ls_inverse);
}

// scale: the scene's extent over the unit extent of scene 2 (positions,
// ranges and the shadow frustum's depth range scale with it)
static void make_point_light(Rng &r, int sm, vct_light &L, double scale = 1.0) {
    std::memset(&L, 0, sizeof(L));
    const float pos[3] = {(float)(scale * r.uni(-0.8, 0.8)), (float)(scale * r.uni(-0.8, 0.8)),
                          (float)(scale * r.uni(-0.8, 0.8))};
    std::memcpy(L.position, pos, 12);
    L.direction[2] = -1.0f;
    for (int k = 0; k < 3; ++k) L.color[k] = (float)r.uni(0.5, 1.0);
    L.intensity = 1.0f;
    L.range = (float)(scale * r.uni(0.8, 1.6));
    L.type = VCT_LIGHT_POINT;
    L.enabled = 1;
    L.shadow_caster = 1;
    L.shadow_size = (uint32_t)sm;
    L.perspective = 1;
    // one 120-degree perspective shadow map aimed at the scene centre
    float lp[16], lv[16];
    perspective(2.0943951f, 1.0f, (float)(0.02 * scale), (float)(4.0 * scale), lp);
    float tgt[3] = {0.0f, 0.0f, 0.0f};
    const float d2 = (pos[0] * pos[0] + pos[1] * pos[1]) + pos[2] * pos[2];
    if (d2 < (float)(0.0025 * scale * scale)) tgt[1] = pos[1] - 1.0f;
    float up[3] = {0.0f, 1.0f, 0.0f};
    const float f[3] = {tgt[0] - pos[0], tgt[1] - pos[1], tgt[2] - pos[2]};
    if (std::fabs(f[1]) > 0.99f * std::sqrt((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2])) {
        up[1] = 0.0f;
        up[2] = 1.0f;
    }
    lookat(pos, tgt, up, lv);
    mul4(lp, lv, L.ls);
    inverse4(L.ls, L.ls_inverse);
}

static bool inside_any_box(const Mesh &m, const std::vector<std::array<double, 6>> &boxes, const double p[3]) {
    (void)m;
    for (auto &b : boxes) {
        bool in = true;
        for (int k = 0; k < 3; ++k)
            if (std::fabs(p[k] - b[k]) > b[3 + k] + 0.05) in = false;
        if (in) return true;
    }
    return false;
}

static void make_camera(Rng &r, const std::vector<std::array<double, 6>> &boxes, const Mesh &m, double range, int W,
                        int H, float znear, vct_camera &cam, float zfar = 10.0f) {
    double eye[3];
    for (int tries = 0; tries < 1000; ++tries) {
        for (int k = 0; k < 3; ++k) eye[k] = r.uni(-range, range);
        if (!inside_any_box(m, boxes, eye)) break;
    }
    const double yaw = r.uni(0.0, 6.283185307179586), pitch = r.uni(-0.3, 0.3);
    const float e[3] = {(float)eye[0], (float)eye[1], (float)eye[2]};
    const float front[3] = {(float)(std::cos(pitch) * std::cos(yaw)), (float)std::sin(pitch),
                            (float)(std::cos(pitch) * std::sin(yaw))};
    const float tgt[3] = {e[0] + front[0], e[1] + front[1], e[2] + front[2]};
    const float up[3] = {0.0f, 1.0f, 0.0f};
    std::memset(&cam, 0, sizeof(cam));
    cam.width = W;
    cam.height = H;
    std::memcpy(cam.eye, e, 12);
    lookat(e, tgt, up, cam.view);
    // explicit 60-degree vertical field of view (the reference passes 45 as
    // radians to glm::perspective, src/Camera.h:18 + Application.cpp:200)
    perspective(1.0471975512f, (float)W / (float)H, znear, zfar, cam.projection);
    cam.cull_back = 1;
}

extern "C" {

int vct_app_default_config(int scene, vct_app_config *o) {
    if (!o || (scene != 1 && scene != 2 && scene != 5)) return VCT_ERR_INVALID;
    std::memset(o, 0, sizeof(*o));
    o->scene = scene;
    o->seed = (uint64_t)scene;
    o->storage = VCT_STORE_FP16;
    o->lighting = 1;
    o->msaa = 1;
    o->fill_holes = 0;
    o->device = 0;
    o->temporal_filter = 0;
    o->temporal_decay = 0.8f;
    o->track_camera = 0;
    o->warp_texture = 0;
    if (scene == 1) {
        o->dim = 64;
        o->levels = 7;
        o->anisotropic = 0;
        o->width = o->height = 256;
        o->shadow_size = 512;
        o->reflections = 0;
    } else if (scene == 2) {
        o->dim = 256;
        o->levels = 8;
        o->anisotropic = 1;
        o->width = 1920;
        o->height = 1080;
        o->shadow_size = 1024;
        o->reflections = 1;
    } else {  // the warped 4K workload: FP8 coarse levels, warp texture at high resolution 4
        o->dim = 256;
        o->levels = 8;
        o->anisotropic = 1;
        o->storage = VCT_STORE_FP8;
        o->width = 3840;
        o->height = 2160;
        o->shadow_size = 1024;
        o->reflections = 1;
        o->warp_texture = 1;
        o->warp_high_resolution = 4.0f;
    }
    return VCT_OK;
}

static void fill_defaults(vct_app_config &c) {
    vct_app_config d;
    vct_app_default_config(c.scene, &d);
    if (!c.seed) c.seed = d.seed;
    if (c.dim <= 0) c.dim = d.dim;
    if (c.levels <= 0) c.levels = d.levels;
    if (c.storage < 0) c.storage = d.storage;
    if (c.anisotropic < 0) c.anisotropic = d.anisotropic;
    if (c.width <= 0) c.width = d.width;
    if (c.height <= 0) c.height = d.height;
    if (c.shadow_size <= 0) c.shadow_size = d.shadow_size;
    if (c.lighting < 0) c.lighting = d.lighting;
    if (c.msaa < 0) c.msaa = d.msaa;
    if (c.reflections < 0) c.reflections = d.reflections;
    if (c.fill_holes < 0) c.fill_holes = d.fill_holes;
    if (!(c.warp_high_resolution > 0.0f)) c.warp_high_resolution = 2.0f;  // Settings (Application.h:87)
}

// voxelTrackCamera (Application.cpp:187-191): the grid centre snaps to the
// eye on cells of 2^levels voxels, so that the coarsest level's texels keep
// their places between snaps.  Returns whether the centre changed.
static bool track_camera(vct_app *a) {
    bool changed = false;
    for (int k = 0; k < 3; ++k) {
        const float cell = std::pow(2.0f, (float)a->cfg.levels) * (a->extent_max[k] - a->extent_min[k]) /
                           (float)a->cfg.dim;
        const float c = std::floor(a->cam.eye[k] / cell) * cell;
        changed = changed || c != a->center[k];
        a->center[k] = c;
    }
    return changed;
}

static int app_finish(vct_app *a, vct_app **out);

int vct_app_create(vct_app **out, const vct_app_config *cfg_in) {
    if (!out || !cfg_in) return VCT_ERR_INVALID;
    *out = nullptr;
    if (cfg_in->scene != 1 && cfg_in->scene != 2 && cfg_in->scene != 5) return VCT_ERR_INVALID;
    vct_app *a = new vct_app();
    a->cfg = *cfg_in;
    fill_defaults(a->cfg);
    const vct_app_config &c = a->cfg;
    Rng r(c.seed * 0x9E3779B97F4A7C15ull + 12345);
    std::vector<std::array<double, 6>> boxes;
    std::vector<double> centers;
    double bound;
    if (c.scene == 1) {
        // room [-1,1]^3 (12 inward triangles) + 2 boxes (centres U[-0.6,0.6]^3,
        // half-extents U[0.1,0.3])
        const double room[3] = {1, 1, 1};
        add_box(a->mesh, room, true, 0, r);
        centers.insert(centers.end(), {0, 0, 0});
        for (int b = 0; b < 2; ++b) {
            double cen[3], h[3];
            for (int k = 0; k < 3; ++k) cen[k] = r.uni(-0.6, 0.6);
            for (int k = 0; k < 3; ++k) h[k] = r.uni(0.1, 0.3);
            add_box(a->mesh, h, false, (uint32_t)(b + 1), r);
            centers.insert(centers.end(), {cen[0], cen[1], cen[2]});
            boxes.push_back({cen[0], cen[1], cen[2], h[0], h[1], h[2]});
        }
        a->mesh.nobj = 3;
        // voxel extent +-(1 + 1/(dim-1)): the room walls fall on voxel centres
        const float e = (float)(1.0 + 1.0 / (double)(c.dim - 1));
        for (int k = 0; k < 3; ++k) {
            a->extent_min[k] = -e;
            a->extent_max[k] = e;
            a->center[k] = 0.0f;
        }
        bound = 1.8;
    } else {
        // scene 2: 2000 boxes (half-extents lognormal(mu=-3, sigma=0.5)), 64
        // icospheres (subdivision 4, radius U[0.02,0.1]) in [-1,1]^3.  Scene 5:
        // 8000 boxes and 512 icospheres in [-4,4]^3, positions scaled by 4 and
        // sizes by 2.5 -- the cube root of the volume each object has to
        // itself (64 x the volume for 4.12 x the objects), so the objects fill
        // the same quarter of the volume as in scene 2 (sizes x4 would fill it
        // completely).  Centres are drawn from 0.9 of the extent, so that
        // objects of the largest sizes start inside the grid.  Each object
        // moves rigidly.
        const double S = c.scene == 5 ? 4.0 : 1.0, Z = c.scene == 5 ? 2.5 : 1.0;
        const int nbox = c.scene == 5 ? 8000 : 2000, nsph = c.scene == 5 ? 512 : 64;
        for (int b = 0; b < nbox; ++b) {
            double h[3];
            for (int k = 0; k < 3; ++k) h[k] = Z * std::exp(-3.0 + 0.5 * r.normal());
            add_box(a->mesh, h, false, (uint32_t)b, r);
            double cen[3];
            for (int k = 0; k < 3; ++k) cen[k] = r.uni(-0.9 * S, 0.9 * S);
            centers.insert(centers.end(), {cen[0], cen[1], cen[2]});
            boxes.push_back({cen[0], cen[1], cen[2], h[0], h[1], h[2]});
        }
        for (int s = 0; s < nsph; ++s) {
            double rad = r.uni(0.02 * Z, 0.1 * Z);
            add_icosphere(a->mesh, rad, 4, (uint32_t)(nbox + s), r);
            double cen[3];
            for (int k = 0; k < 3; ++k) cen[k] = r.uni(-0.9 * S, 0.9 * S);
            centers.insert(centers.end(), {cen[0], cen[1], cen[2]});
            boxes.push_back({cen[0], cen[1], cen[2], rad, rad, rad});
        }
        a->mesh.nobj = nbox + nsph;
        a->mesh.rough.resize(a->mesh.nobj);
        for (auto &v : a->mesh.rough) v = (float)r.uni(0.1, 0.8);
        for (int k = 0; k < 3; ++k) {
            a->extent_min[k] = (float)-S;
            a->extent_max[k] = (float)S;
            a->center[k] = 0.0f;
        }
        bound = 1.8 * S;
        a->dynamic = true;
    }
    // per-object motion (static scenes keep omega = vel = 0)
    a->motion.resize(a->mesh.nobj);
    for (int o = 0; o < a->mesh.nobj; ++o) {
        ObjMotion &mo = a->motion[o];
        for (int k = 0; k < 3; ++k) mo.c0[k] = centers[(size_t)o * 3 + k];
        double ax[3] = {r.normal(), r.normal(), r.normal()};
        double l = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
        if (l < 1e-12) {
            ax[0] = 1;
            l = 1;
        }
        for (int k = 0; k < 3; ++k) mo.axis[k] = ax[k] / l;
        if (a->dynamic) {
            const double S = c.scene == 5 ? 4.0 : 1.0;  // translation scales with the extent
            mo.omega = r.uni(0.0, 5.0 * 3.141592653589793 / 180.0);
            for (int k = 0; k < 3; ++k) mo.vel[k] = r.uni(-0.01 * S, 0.01 * S);
        } else {
            mo.omega = 0.0;
            for (int k = 0; k < 3; ++k) mo.vel[k] = 0.0;
        }
    }
    // lights
    if (c.scene == 1) {
        a->lights.resize(1);
        make_dir_light(r, bound, c.shadow_size, a->lights[0]);
    } else if (c.scene == 2) {
        a->lights.resize(5);
        make_dir_light(r, bound, c.shadow_size, a->lights[0]);
        for (int i = 1; i < 5; ++i) make_point_light(r, c.shadow_size, a->lights[i]);
    } else {  // 8 lights: 1 directional + 7 point lights, each with its shadow map
        a->lights.resize(8);
        make_dir_light(r, bound, c.shadow_size, a->lights[0]);
        for (int i = 1; i < 8; ++i) make_point_light(r, c.shadow_size, a->lights[i], 4.0);
    }
    a->boxes = boxes;
    a->cam_range = c.scene == 5 ? 2.4 : (c.scene == 1 ? 0.5 : 0.6);
    a->cam_znear = c.scene == 5 ? 0.08f : (c.scene == 1 ? 0.05f : 0.02f);
    a->cam_zfar = c.scene == 5 ? 40.0f : 10.0f;
    make_camera(r, boxes, a->mesh, a->cam_range, c.width, c.height, a->cam_znear, a->cam, a->cam_zfar);
    return app_finish(a, out);
}

// trace settings, the camera snap, the GPU context and the scene upload,
// common to every way of building the scene
static int app_finish(vct_app *a, vct_app **out) {
    const vct_app_config &c = a->cfg;
    // trace settings: VCTSettings defaults (Application.h:96-98)
    vct_trace_params &t = a->trace;
    std::memset(&t, 0, sizeof(t));
    t.diffuse = {16, 1.0471975512f, 1.0f, 1.0f, 0.5f};
    t.specular = {32, 0.5235987756f, 1.7f, 0.5f, 0.1f};
    t.enable_reflections = c.reflections;
    t.cone_angle_from_roughness = 1;
    t.draw_occlusion = 1;
    t.ambient_scale = 1.0f;
    t.reflect_scale = 1.0f;
    std::memcpy(t.eye, a->cam.eye, 12);
    t.warp_texture = c.warp_texture;
    if (c.track_camera) track_camera(a);
    // GPU context
    vct_grid_desc d;
    std::memset(&d, 0, sizeof(d));
    d.dim = c.dim;
    d.levels = c.levels;
    std::memcpy(d.center, a->center, 12);
    std::memcpy(d.vmin, a->extent_min, 12);
    std::memcpy(d.vmax, a->extent_max, 12);
    d.storage = c.storage;
    d.anisotropic = c.anisotropic;
    d.device = c.device;
    a->model.resize((size_t)a->mesh.nobj * 16);
    if (c.device < 0) {  // scene and uniforms only, no GPU context
        *out = a;
        return VCT_OK;
    }
    int rc = vct_create(&a->ctx, &d);
    if (rc) {
        delete a;
        return rc;
    }
    vct_mesh m;
    m.num_tris = a->mesh.ntri();
    m.num_objects = a->mesh.nobj;
    m.positions = a->mesh.pos.data();
    m.normals = a->mesh.nrm.data();
    m.tangents = a->mesh.tan.data();
    m.albedo = a->mesh.albedo.data();
    m.object = a->mesh.obj.data();
    m.roughness = a->mesh.rough.empty() ? nullptr : a->mesh.rough.data();
    m.vertex_tangents = a->mesh.vtan.empty() ? nullptr : a->mesh.vtan.data();
    m.texcoords = a->mesh.uv.empty() ? nullptr : a->mesh.uv.data();
    m.texture = a->mesh.texid.empty() ? nullptr : a->mesh.texid.data();
    if ((rc = vct_set_scene(a->ctx, &m)) ||
        (rc = vct_set_textures(a->ctx, a->textures.empty() ? nullptr : a->textures.data(), (int)a->textures.size())) ||
        (rc = vct_set_lights(a->ctx, a->lights.data(), (int)a->lights.size()))) {
        vct_app_destroy(a);
        return rc;
    }
    *out = a;
    return VCT_OK;
}

int vct_app_create_from_obj(vct_app **out, const vct_app_config *cfg_in, const char *path) {
    if (!out || !cfg_in || !path) return VCT_ERR_INVALID;
    *out = nullptr;
    vct_obj *obj = nullptr;
    char err[256] = {0};
    int rc = vct_obj_load(path, &obj, err, (int)sizeof(err));
    if (rc) return rc;
    vct_app *a = new vct_app();
    a->cfg = *cfg_in;
    a->cfg.scene = 1;  // defaults of the small static scene for unset fields
    fill_defaults(a->cfg);
    const vct_app_config &c = a->cfg;
    vct_mesh m;
    vct_obj_mesh(obj, &m);
    a->mesh.pos.assign(m.positions, m.positions + (size_t)m.num_tris * 9);
    a->mesh.nrm.assign(m.normals, m.normals + (size_t)m.num_tris * 9);
    a->mesh.tan.assign(m.tangents, m.tangents + (size_t)m.num_tris * 3);
    if (m.vertex_tangents) a->mesh.vtan.assign(m.vertex_tangents, m.vertex_tangents + (size_t)m.num_tris * 9);
    a->mesh.albedo.assign(m.albedo, m.albedo + m.num_tris);
    a->mesh.obj.assign(m.object, m.object + m.num_tris);
    if (m.roughness) a->mesh.rough.assign(m.roughness, m.roughness + 1);
    a->mesh.nobj = 1;
    const vct_texture *tex = nullptr;
    int ntex = 0;
    vct_obj_textures(obj, &tex, &ntex);
    if (ntex > 0) {  // the diffuse maps, copied: the obj is freed below
        a->mesh.uv.assign(m.texcoords, m.texcoords + (size_t)m.num_tris * 6);
        a->mesh.texid.assign(m.texture, m.texture + m.num_tris);
        for (int i = 0; i < ntex; ++i) {
            size_t n = 0;
            for (int l = 0; l < tex[i].levels; ++l)
                n += (size_t)std::max(1, tex[i].width >> l) * (size_t)std::max(1, tex[i].height >> l);
            a->tex_texels.emplace_back(tex[i].texels, tex[i].texels + n);
            a->textures.push_back(tex[i]);
        }
        for (int i = 0; i < ntex; ++i) a->textures[i].texels = a->tex_texels[i].data();
    }
    float bmin[3], bmax[3];
    vct_obj_bounds(obj, bmin, bmax, nullptr);
    vct_obj_free(obj);
    float half = 0.0f, cen[3];
    for (int k = 0; k < 3; ++k) {
        cen[k] = 0.5f * (bmin[k] + bmax[k]);
        half = std::max(half, 0.5f * (bmax[k] - bmin[k]));
    }
    half = half > 0.0f ? 1.05f * half : 1.0f;
    for (int k = 0; k < 3; ++k) {
        a->center[k] = cen[k];
        a->extent_min[k] = -half;
        a->extent_max[k] = half;
    }
    a->motion.resize(1);
    std::memset(&a->motion[0], 0, sizeof(ObjMotion));
    a->motion[0].axis[1] = 1.0;
    a->dynamic = false;
    // one directional light from above, its shadow frustum around the bounds
    a->lights.resize(1);
    vct_light &L = a->lights[0];
    std::memset(&L, 0, sizeof(L));
    const float dir[3] = {-0.3f, -0.9f, -0.3f};
    const float dl = std::sqrt((dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2]);
    for (int k = 0; k < 3; ++k) L.direction[k] = dir[k] / dl;
    const float dist = 2.5f * half;
    for (int k = 0; k < 3; ++k) L.position[k] = cen[k] - L.direction[k] * dist;
    L.color[0] = L.color[1] = L.color[2] = 1.0f;
    L.intensity = 1.0f;
    L.range = 5.0f * half;
    L.type = VCT_LIGHT_DIRECTIONAL;
    L.enabled = 1;
    L.shadow_caster = 1;
    L.shadow_size = (uint32_t)c.shadow_size;
    float lp[16], lv[16];
    ortho(-1.5f * half, 1.5f * half, -1.5f * half, 1.5f * half, 0.0f, 2.0f * dist, lp);
    const float tgt[3] = {L.position[0] + L.direction[0], L.position[1] + L.direction[1],
                          L.position[2] + L.direction[2]};
    const float up[3] = {0.0f, 1.0f, 0.0f};
    lookat(L.position, tgt, up, lv);
    mul4(lp, lv, L.ls);
    inverse4(L.ls, L.ls_inverse);
    // camera outside the bounds, looking at the centre
    std::memset(&a->cam, 0, sizeof(a->cam));
    a->cam.width = c.width;
    a->cam.height = c.height;
    const float eye[3] = {cen[0] + 1.2f * half, cen[1] + 0.6f * half, cen[2] + 1.8f * half};
    std::memcpy(a->cam.eye, eye, 12);
    lookat(eye, cen, up, a->cam.view);
    perspective(1.0471975512f, (float)c.width / (float)c.height, 0.02f * half, 10.0f * half, a->cam.projection);
    a->cam.cull_back = 1;
    return app_finish(a, out);
}

void vct_app_destroy(vct_app *a) {
    if (!a) return;
    if (a->ctx) vct_destroy(a->ctx);
    delete a;
}

const char *vct_app_last_error(const vct_app *a) {
    if (!a) return vct_last_error(nullptr);
    if (!a->err.empty()) return a->err.c_str();
    return vct_last_error(a->ctx);
}

vct_ctx *vct_app_ctx(vct_app *a) { return a ? a->ctx : nullptr; }

int vct_app_get_config(const vct_app *a, vct_app_config *out) {
    if (!a || !out) return VCT_ERR_INVALID;
    *out = a->cfg;
    return VCT_OK;
}

int vct_app_scene_sizes(const vct_app *a, int *nt, int *no, int *nl) {
    if (!a) return VCT_ERR_INVALID;
    if (nt) *nt = a->mesh.ntri();
    if (no) *no = a->mesh.nobj;
    if (nl) *nl = (int)a->lights.size();
    return VCT_OK;
}

int vct_app_get_mesh(const vct_app *a, float *pos, float *nrm, float *tan, uint32_t *albedo, uint32_t *obj,
                     float *rough) {
    if (!a) return VCT_ERR_INVALID;
    const Mesh &m = a->mesh;
    if (pos) std::memcpy(pos, m.pos.data(), m.pos.size() * 4);
    if (nrm) std::memcpy(nrm, m.nrm.data(), m.nrm.size() * 4);
    if (tan) std::memcpy(tan, m.tan.data(), m.tan.size() * 4);
    if (albedo) std::memcpy(albedo, m.albedo.data(), m.albedo.size() * 4);
    if (obj) std::memcpy(obj, m.obj.data(), m.obj.size() * 4);
    if (rough && !m.rough.empty()) std::memcpy(rough, m.rough.data(), m.rough.size() * 4);
    return m.rough.empty() ? 0 : 1;
}

int vct_app_get_vertex_tangents(const vct_app *a, float *vtan) {
    if (!a) return VCT_ERR_INVALID;
    if (a->mesh.vtan.empty()) return 0;
    if (vtan) std::memcpy(vtan, a->mesh.vtan.data(), a->mesh.vtan.size() * 4);
    return 1;
}

static void frame_models(const vct_app *a, int frame, float *out) {
    for (int o = 0; o < a->mesh.nobj; ++o) {
        const ObjMotion &mo = a->motion[o];
        double c[3];
        // centres bounce inside 0.95 of the extent (0.95 in scene 2, 3.8 in scene 5)
        const double B = 0.95 * (double)a->extent_max[0];
        for (int k = 0; k < 3; ++k) c[k] = a->dynamic ? reflect_into(mo.c0[k] + mo.vel[k] * frame, B) : mo.c0[k];
        rigid(c, mo.axis, mo.omega * frame, out + (size_t)o * 16);
    }
}

int vct_app_frame_state(vct_app *a, int frame, float *model, vct_light *lights, vct_voxelize_params *vox,
                        vct_inject_params *inj, vct_mip_params *mip, vct_camera *cam, vct_trace_params *trace) {
    if (!a) return VCT_ERR_INVALID;
    if (model) frame_models(a, frame, model);
    if (lights) std::memcpy(lights, a->lights.data(), a->lights.size() * sizeof(vct_light));
    if (vox) {
        std::memset(vox, 0, sizeof(*vox));
        voxel_projections(a->center, a->extent_min, a->extent_max, vox->mvp_x, vox->mvp_y, vox->mvp_z);
        vox->axis_override = -1;
        vox->lighting = a->cfg.lighting;
        vox->atomic_max = 1;
        vox->msaa = a->cfg.msaa;
        std::memcpy(vox->eye, a->cam.eye, 12);
        vox->warp_texture = a->cfg.warp_texture;
    }
    if (inj) {
        inj->voxel_set_opacity = 0.5f;
        inj->light_mask = 0xffffffffu;
        inj->temporal_filter = a->cfg.temporal_filter;
        inj->temporal_decay = a->cfg.temporal_decay;
        inj->warp_texture = a->cfg.warp_texture;
    }
    if (mip) {
        mip->kernel_mode = 0;
        mip->albedo_chain = 1;  // the mipmap pass filters both chains (Application.cpp:877-919)
    }
    if (cam) *cam = a->cam;
    if (trace) *trace = a->trace;
    return VCT_OK;
}

// normal matrix (cofactor / determinant of the upper 3x3), as the pass ABI derives it
static void normal_matrix3(const float *m, float *out) {
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
    const float v[9] = {c00, c10, c20, c01, c11, c21, c02, c12, c22};
    for (int i = 0; i < 9; ++i) out[i] = v[i] / det;
}

int vct_app_surface_points(vct_app *a, int frame, int n, uint64_t seed, vct_gbuffer_texel *out) {
    if (!a || !out || n <= 0) return VCT_ERR_INVALID;
    const Mesh &m = a->mesh;
    const int T = m.ntri();
    if (T == 0) return VCT_ERR_INVALID;
    std::vector<float> model((size_t)m.nobj * 16), nmat((size_t)m.nobj * 9);
    frame_models(a, frame, model.data());
    for (int o = 0; o < m.nobj; ++o) normal_matrix3(model.data() + (size_t)o * 16, nmat.data() + (size_t)o * 9);
    auto world = [&](int t, int v, float w[3]) {
        const float *M = model.data() + (size_t)m.obj[t] * 16;
        const float *p = m.pos.data() + (size_t)t * 9 + v * 3;
        for (int r = 0; r < 3; ++r) {
            float s = M[0 * 4 + r] * p[0];
            s = s + M[1 * 4 + r] * p[1];
            s = s + M[2 * 4 + r] * p[2];
            s = s + M[3 * 4 + r] * 1.0f;
            w[r] = s;
        }
    };
    auto m3 = [](const float *N, const float *v, float o[3]) {
        for (int r = 0; r < 3; ++r) o[r] = (N[0 * 3 + r] * v[0] + N[1 * 3 + r] * v[1]) + N[2 * 3 + r] * v[2];
    };
    std::vector<double> cdf((size_t)T);
    double acc = 0.0;
    for (int t = 0; t < T; ++t) {
        float W[3][3];
        for (int v = 0; v < 3; ++v) world(t, v, W[v]);
        double e1[3], e2[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = W[1][k] - W[0][k];
            e2[k] = W[2][k] - W[0][k];
        }
        const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                     cz = e1[0] * e2[1] - e1[1] * e2[0];
        acc += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
        cdf[(size_t)t] = acc;
    }
    Rng r(seed * 0x9E3779B97F4A7C15ull + 777);
    for (int i = 0; i < n; ++i) {
        const double u = r.uni() * acc;
        int t = (int)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        if (t >= T) t = T - 1;
        double s1 = std::sqrt(r.uni()), s2 = r.uni();
        float b[3] = {(float)(1.0 - s1), (float)(s1 * (1.0 - s2)), (float)(s1 * s2)};
        float W[3][3], Nv[3][3], Tv[3][3], Bv[3][3], Tw[3];
        const float *Nm = nmat.data() + (size_t)m.obj[t] * 9;
        m3(Nm, m.tan.data() + (size_t)t * 3, Tw);
        for (int v = 0; v < 3; ++v) {
            world(t, v, W[v]);
            m3(Nm, m.nrm.data() + (size_t)t * 9 + v * 3, Nv[v]);
            if (!m.vtan.empty()) m3(Nm, m.vtan.data() + (size_t)t * 9 + v * 3, Tw);
            const float dtn = (Tw[0] * Nv[v][0] + Tw[1] * Nv[v][1]) + Tw[2] * Nv[v][2];
            float tt[3] = {Tw[0] - dtn * Nv[v][0], Tw[1] - dtn * Nv[v][1], Tw[2] - dtn * Nv[v][2]};
            nrm3(tt, Tv[v]);
            cross3(Nv[v], Tv[v], Bv[v]);
        }
        vct_gbuffer_texel &g = out[i];
        for (int k = 0; k < 3; ++k) {
            g.position[k] = (b[0] * W[0][k] + b[1] * W[1][k]) + b[2] * W[2][k];
            g.normal[k] = (b[0] * Nv[0][k] + b[1] * Nv[1][k]) + b[2] * Nv[2][k];
            g.tangent[k] = (b[0] * Tv[0][k] + b[1] * Tv[1][k]) + b[2] * Tv[2][k];
            g.bitangent[k] = (b[0] * Bv[0][k] + b[1] * Bv[1][k]) + b[2] * Bv[2][k];
        }
        g.albedo = m.albedo[(size_t)t];
        g.roughness = m.rough.empty() ? 0.5f : m.rough[m.obj[t]];
        g.covered = 1;
        g.triangle = (uint32_t)t;
    }
    return VCT_OK;
}

int vct_app_camera(const vct_app *a, uint64_t seed, int width, int height, vct_camera *out) {
    if (!a || !out || width < 1 || height < 1) return VCT_ERR_INVALID;
    if (a->cam_range <= 0.0) return VCT_ERR_STATE;  // an OBJ scene: one fitted camera only
    Rng r(seed * 0x9E3779B97F4A7C15ull + 4242);
    make_camera(r, a->boxes, a->mesh, a->cam_range, width, height, a->cam_znear, *out, a->cam_zfar);
    return VCT_OK;
}

int vct_app_build_radiance(vct_app *a, int frame) {
    if (!a) return VCT_ERR_INVALID;
    if (!a->ctx) return app_fail(a, VCT_ERR_STATE, "vct_app_build_radiance: app was created without a device");
    if (a->cfg.track_camera && track_camera(a)) {
        const int rc = vct_set_extent(a->ctx, a->center, a->extent_min, a->extent_max);
        if (rc) return rc;
    }
    vct_voxelize_params vox;
    vct_inject_params inj;
    vct_mip_params mip;
    vct_app_frame_state(a, frame, a->model.data(), nullptr, &vox, &inj, &mip, nullptr, nullptr);
    int rc;
    if ((rc = vct_set_transforms(a->ctx, a->model.data(), a->mesh.nobj))) return rc;
    if ((rc = vct_shadowmaps(a->ctx))) return rc;  // every caster's map (Application.cpp:212-233 per light)
    if (a->cfg.warp_texture) {  // the occupancy grid and warp texture of this frame (Application.cpp:235-575)
        vct_warpmap_params wp;
        vct_app_warpmap_params(a, &wp);
        if ((rc = vct_warpmap(a->ctx, &vox, &wp))) return rc;
    }
    if ((rc = a->cfg.tessellation ? vct_voxelize_tessellated(a->ctx, &vox) : vct_voxelize(a->ctx, &vox))) return rc;
    if ((rc = vct_normalize_inject(a->ctx, &inj))) return rc;
    if (a->cfg.fill_holes && (rc = vct_fill_holes(a->ctx))) return rc;
    if ((rc = vct_mip(a->ctx, &mip))) return rc;
    return VCT_OK;
}

int vct_app_warpmap_params(const vct_app *a, vct_warpmap_params *out) {
    if (!a || !out) return VCT_ERR_INVALID;
    out->high_resolution = a->cfg.warp_high_resolution;  // Settings (src/Application.h:80-92), 2.0 by default
    out->low_resolution = 0.5f;
    out->axes[0] = out->axes[1] = out->axes[2] = 1;
    out->linear = 0;
    out->use_weights_texture = 1;
    return VCT_OK;
}

int vct_app_compose_params(vct_app *a, int frame, vct_compose_params *out) {
    (void)frame;  // the camera is static in every scene
    if (!a || !out) return VCT_ERR_INVALID;
    vct_compose_params p;
    std::memset(&p, 0, sizeof(p));
    p.cook_torrance = 1;
    p.enable_diffuse = 1;
    p.enable_specular = 1;
    p.enable_indirect = 1;
    p.enable_reflections = a->trace.enable_reflections;
    p.postprocess = 1;
    p.ambient_scale = 1.0f;
    p.reflect_scale = 1.0f;
    p.shininess = 32.0f;  // Material() default (src/Graphics/Mesh.h:29)
    std::memcpy(p.eye, a->cam.eye, 12);
    *out = p;
    return VCT_OK;
}

int vct_app_render(vct_app *a, int frame) {
    if (!a) return VCT_ERR_INVALID;
    int rc = vct_app_build_radiance(a, frame);
    if (rc) return rc;
    if ((rc = vct_gbuffer(a->ctx, &a->cam))) return rc;
    if ((rc = vct_trace(a->ctx, &a->trace))) return rc;
    vct_compose_params cp;
    vct_app_compose_params(a, frame, &cp);
    return vct_compose(a->ctx, &cp);
}

void vct_math_ortho(float l, float r, float b, float t, float n, float f, float out[16]) { ortho(l, r, b, t, n, f, out); }
void vct_math_lookat(const float eye[3], const float center[3], const float up[3], float out[16]) {
    lookat(eye, center, up, out);
}
void vct_math_perspective(float fovy, float aspect, float n, float f, float out[16]) {
    perspective(fovy, aspect, n, f, out);
}
int vct_math_inverse4(const float m[16], float out[16]) { return inverse4(m, out) ? 1 : 0; }
void vct_math_mul4(const float a[16], const float b[16], float out[16]) { mul4(a, b, out); }
void vct_math_voxel_projections(const float c[3], const float mn[3], const float mx[3], float mvp_x[16],
                                float mvp_y[16], float mvp_z[16]) {
    voxel_projections(c, mn, mx, mvp_x, mvp_y, mvp_z);
}

}  // extern "C"
