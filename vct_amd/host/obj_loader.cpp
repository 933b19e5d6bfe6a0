// obj_loader.cpp -- Wavefront OBJ/MTL ingestion into the triangle soup the
// passes take (vct_mesh).  Restates what Mesh::loadMesh (src/Graphics/
// Mesh.cpp:42-266, on tinyobjloader) gives the renderer:
//   * faces triangulated, vertex attributes by (v, vn, vt) index triple;
//     1-based indices, negative ones relative to the end of the list so far;
//   * texture v flipped to 1 - v (Mesh.cpp:150);
//   * the per-face tangent of learnopengl's normal-mapping rule from the
//     position and uv deltas (Mesh.cpp:174-189), normalized, and the
//     per-vertex tangents the reference draws with: vertices are shared by
//     their (v, vt, vn) index triple (Mesh.cpp:138-167), every face adds its
//     unnormalized tangent to its three vertices, and each vertex's sum is
//     normalized (Mesh.cpp:190-197);
//   * the diffuse colour Kd of the face's material (usemtl/newmtl) as RGBA8,
//     and the PBR roughness Pr per material;
//   * the material's diffuse map (map_Kd, path relative to the .obj with
//     backslashes made slashes, Mesh.cpp:62-81) when it is a .dds or .png
//     file: the faces of that material carry their corners' uvs and the map's
//     index (maps are loaded once per path); a corner without vt has uv (0, 0).
// Differences, on purpose: polygons become triangle fans (tinyobjloader's
// default triangulation for convex faces); a face without vn gets its face
// normal and one without usable uvs an orthonormal tangent from the normal
// (the reference would produce NaN there) and adds nothing to its vertices'
// tangent sums (a vertex whose sum is zero keeps its face's tangent); maps in
// other formats (stb_image's JPEG, TGA, BMP ...) and the other maps (map_Pr,
// map_Ks, ...) are not decoded -- only DDS and PNG decoders are built -- so such a
// face's albedo is its material's Kd, white without a material, and its
// roughness Pr (0.5 if the material has none); a map that cannot be read
// is skipped the same way (the reference logs it and draws on).  One object:
// the file is one actor, as in Scene.
#include <algorithm>
#include <cctype>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "vct_app.h"

struct vct_obj {
    std::vector<float> pos, nrm, tan;
    std::vector<float> vtan;  // [ntri][3][3]: the normalized per-vertex tangent sums
    std::vector<float> uv;    // [ntri][3][2]
    std::vector<int32_t> texid;
    std::vector<vct_image *> images;
    std::vector<vct_texture> textures;
    std::vector<uint32_t> albedo, object;
    std::vector<float> rough;  // one value: the file is one object
    int nmaterials = 0;
    float bmin[3], bmax[3];
    ~vct_obj() {
        for (vct_image *i : images) vct_dds_free(i);
    }
};

namespace {

struct Material {
    float kd[3] = {1.0f, 1.0f, 1.0f};
    float pr = -1.0f;  // no roughness given
    std::string map_kd;
    int tex = -2;  // index into vct_obj::textures; -1: no usable map; -2: not looked up yet
};

uint32_t pack_rgba8(const float c[3]) {
    uint32_t v = 0;
    for (int k = 0; k < 3; ++k) {
        const float x = std::min(std::max(c[k], 0.0f), 1.0f);
        v |= ((uint32_t)std::lrint(x * 255.0f) & 0xffu) << (8 * k);
    }
    return v | 0xff000000u;
}

void set_err(char *err, int len, const std::string &msg) {
    if (err && len > 0) std::snprintf(err, (size_t)len, "%s", msg.c_str());
}

// resolves a 1-based or negative OBJ index against n entries; -1 if absent
int resolve(const std::string &tok, int n, bool *bad) {
    if (tok.empty()) return -1;
    char *end = nullptr;
    const long v = std::strtol(tok.c_str(), &end, 10);
    if (*end != '\0' || v == 0) {
        *bad = true;
        return -1;
    }
    const long i = v > 0 ? v - 1 : n + v;
    if (i < 0 || i >= n) {
        *bad = true;
        return -1;
    }
    return (int)i;
}

bool load_mtl(const std::string &path, std::map<std::string, Material> &mats, std::string &why) {
    std::ifstream f(path);
    if (!f) {
        why = "cannot open material library " + path;
        return false;
    }
    std::string line, cur;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::string key;
        if (!(ss >> key) || key[0] == '#') continue;
        if (key == "newmtl") {
            ss >> cur;
            mats[cur] = Material();
        } else if (key == "Kd" && !cur.empty()) {
            Material &m = mats[cur];
            ss >> m.kd[0] >> m.kd[1] >> m.kd[2];
        } else if (key == "Pr" && !cur.empty()) {
            ss >> mats[cur].pr;
        } else if (key == "map_Kd" && !cur.empty()) {
            std::string rest, tok;
            while (ss >> tok) rest = tok;  // the file name ends the statement
            std::replace(rest.begin(), rest.end(), '\\', '/');
            mats[cur].map_kd = rest;
        }
    }
    return true;
}

void normalize3(float v[3]) {
    const float l = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    if (l > 0.0f)
        for (int k = 0; k < 3; ++k) v[k] /= l;
}

}  // namespace

extern "C" {

int vct_obj_load(const char *path, vct_obj **out, char *err, int errlen) {
    if (!path || !out) return VCT_ERR_INVALID;
    *out = nullptr;
    std::ifstream f(path);
    if (!f) {
        set_err(err, errlen, std::string("cannot open ") + path);
        return VCT_ERR_INVALID;
    }
    const std::string sp(path);
    const std::string dir = sp.substr(0, sp.find_last_of('/') + 1);
    std::vector<float> v, vn, vt;
    std::map<std::string, Material> mats;
    Material none;
    none.tex = -1;
    Material *cur = &none;
    std::map<std::string, int> tex_by_path;
    bool any_pr = false;
    float pr = 0.5f;
    vct_obj *o = new vct_obj();
    // shared vertices (Mesh.cpp:138-167): index triple -> slot, and each
    // slot's tangent sum in float, as glm::vec3 accumulates it
    std::map<std::array<int, 3>, int> vmap;
    std::vector<float> vacc;
    std::vector<int> corner_slot;  // [ntri][3]
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        std::istringstream ss(line);
        std::string key;
        if (!(ss >> key) || key[0] == '#') continue;
        if (key == "v") {
            float x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            v.insert(v.end(), {x, y, z});
        } else if (key == "vn") {
            float x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            vn.insert(vn.end(), {x, y, z});
        } else if (key == "vt") {
            float x = 0, y = 0;
            ss >> x >> y;
            vt.insert(vt.end(), {x, 1.0f - y});  // Mesh.cpp:150
        } else if (key == "mtllib") {
            std::string name;
            ss >> name;
            std::string why;
            if (!load_mtl(dir + name, mats, why)) {
                set_err(err, errlen, why);
                delete o;
                return VCT_ERR_INVALID;
            }
        } else if (key == "usemtl") {
            std::string name;
            ss >> name;
            auto it = mats.find(name);
            cur = it == mats.end() ? &none : &it->second;
            if (cur->tex == -2) {  // its diffuse map, loaded once per path
                cur->tex = -1;
                const std::string &mk = cur->map_kd;
                std::string ext = mk.size() > 4 ? mk.substr(mk.size() - 4) : std::string();
                for (char &ch : ext) ch = (char)std::tolower((unsigned char)ch);
                const bool dds = ext == ".dds", png = ext == ".png";  // GLHelper.cpp:166-171
                if (dds || png) {
                    auto tp = tex_by_path.find(mk);
                    if (tp != tex_by_path.end()) {
                        cur->tex = tp->second;
                    } else {
                        vct_image *img = nullptr;
                        const std::string full = dir + mk;
                        const int rc = dds ? vct_dds_load(full.c_str(), &img, nullptr, 0)
                                           : vct_png_load(full.c_str(), &img, nullptr, 0);
                        if (rc == VCT_OK) {
                            vct_texture t;
                            vct_dds_texture(img, &t);
                            o->images.push_back(img);
                            o->textures.push_back(t);
                            cur->tex = (int)o->textures.size() - 1;
                        }
                        tex_by_path[mk] = cur->tex;
                    }
                }
            }
            if (cur->pr >= 0.0f) {
                any_pr = true;
                pr = cur->pr;
            }
        } else if (key == "f") {
            std::vector<int> iv, it_, in;
            std::string tok;
            bool bad = false;
            while (ss >> tok) {
                std::string a, b, c;
                const size_t s1 = tok.find('/');
                if (s1 == std::string::npos) {
                    a = tok;
                } else {
                    a = tok.substr(0, s1);
                    const size_t s2 = tok.find('/', s1 + 1);
                    if (s2 == std::string::npos) {
                        b = tok.substr(s1 + 1);
                    } else {
                        b = tok.substr(s1 + 1, s2 - s1 - 1);
                        c = tok.substr(s2 + 1);
                    }
                }
                iv.push_back(resolve(a, (int)v.size() / 3, &bad));
                it_.push_back(resolve(b, (int)vt.size() / 2, &bad));
                in.push_back(resolve(c, (int)vn.size() / 3, &bad));
                if (iv.back() < 0) bad = true;
            }
            if (bad || iv.size() < 3) {
                set_err(err, errlen, "bad face at line " + std::to_string(lineno));
                delete o;
                return VCT_ERR_INVALID;
            }
            for (size_t k = 1; k + 1 < iv.size(); ++k) {  // fan (0, k, k+1)
                const size_t c3[3] = {0, k, k + 1};
                float P[3][3], N[3][3], T[3][2];
                bool has_n = true, has_t = true;
                for (int j = 0; j < 3; ++j) {
                    const size_t q = c3[j];
                    for (int d = 0; d < 3; ++d) P[j][d] = v[(size_t)iv[q] * 3 + d];
                    if (in[q] >= 0)
                        for (int d = 0; d < 3; ++d) N[j][d] = vn[(size_t)in[q] * 3 + d];
                    else
                        has_n = false;
                    if (it_[q] >= 0)
                        for (int d = 0; d < 2; ++d) T[j][d] = vt[(size_t)it_[q] * 2 + d];
                    else
                        has_t = false;
                }
                const float e1[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
                const float e2[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
                float fn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                               e1[0] * e2[1] - e1[1] * e2[0]};
                normalize3(fn);
                if (!has_n)
                    for (int j = 0; j < 3; ++j) std::memcpy(N[j], fn, 12);
                float tg[3] = {0.0f, 0.0f, 0.0f};
                bool tan_ok = false;
                if (has_t) {  // Mesh.cpp:174-189
                    const float du1 = T[1][0] - T[0][0], dv1 = T[1][1] - T[0][1];
                    const float du2 = T[2][0] - T[0][0], dv2 = T[2][1] - T[0][1];
                    const float det = du1 * dv2 - du2 * dv1;
                    if (std::fabs(det) > 1e-20f) {
                        const float inv = 1.0f / det;
                        for (int d = 0; d < 3; ++d) tg[d] = inv * (dv2 * e1[d] - dv1 * e2[d]);
                        const float l2 = (tg[0] * tg[0] + tg[1] * tg[1]) + tg[2] * tg[2];
                        tan_ok = std::isfinite(l2) && l2 > 0.0f;
                    }
                }
                if (!tan_ok) {  // any unit vector orthogonal to the face normal
                    const float ax[3] = {std::fabs(fn[1]) < 0.9f ? 0.0f : 1.0f, std::fabs(fn[1]) < 0.9f ? 1.0f : 0.0f,
                                         0.0f};
                    tg[0] = ax[1] * fn[2] - ax[2] * fn[1];
                    tg[1] = ax[2] * fn[0] - ax[0] * fn[2];
                    tg[2] = ax[0] * fn[1] - ax[1] * fn[0];
                    if (tg[0] == 0.0f && tg[1] == 0.0f && tg[2] == 0.0f) tg[0] = 1.0f;  // zero-area face: no normal
                }
                for (int j = 0; j < 3; ++j) {
                    const size_t q = c3[j];
                    const std::array<int, 3> key = {iv[q], it_[q], in[q]};
                    auto ins = vmap.emplace(key, (int)(vacc.size() / 3));
                    if (ins.second) vacc.insert(vacc.end(), {0.0f, 0.0f, 0.0f});
                    const int slot = ins.first->second;
                    corner_slot.push_back(slot);
                    if (tan_ok)  // Mesh.cpp:190-193: the unnormalized face tangent
                        for (int d = 0; d < 3; ++d) vacc[(size_t)slot * 3 + d] += tg[d];
                }
                normalize3(tg);
                for (int j = 0; j < 3; ++j) {
                    o->pos.insert(o->pos.end(), P[j], P[j] + 3);
                    o->nrm.insert(o->nrm.end(), N[j], N[j] + 3);
                }
                o->tan.insert(o->tan.end(), tg, tg + 3);
                for (int j = 0; j < 3; ++j) {
                    o->uv.push_back(it_[c3[j]] >= 0 ? T[j][0] : 0.0f);
                    o->uv.push_back(it_[c3[j]] >= 0 ? T[j][1] : 0.0f);
                }
                o->texid.push_back(cur->tex >= 0 ? cur->tex : -1);
                o->albedo.push_back(pack_rgba8(cur->kd));
                o->object.push_back(0u);
            }
        }
        // o, g, s and the other statements do not change the triangle soup
    }
    if (o->albedo.empty()) {
        set_err(err, errlen, std::string("no faces in ") + path);
        delete o;
        return VCT_ERR_INVALID;
    }
    // Mesh.cpp:195-197: every vertex's tangent sum normalized
    const size_t ntri = o->albedo.size();
    o->vtan.resize(ntri * 9);
    for (size_t t = 0; t < ntri; ++t)
        for (int j = 0; j < 3; ++j) {
            const float *a = &vacc[(size_t)corner_slot[t * 3 + j] * 3];
            float tv[3] = {a[0], a[1], a[2]};
            const float l2 = (tv[0] * tv[0] + tv[1] * tv[1]) + tv[2] * tv[2];
            if (l2 > 0.0f && std::isfinite(l2)) normalize3(tv);
            else std::memcpy(tv, &o->tan[t * 3], 12);
            std::memcpy(&o->vtan[t * 9 + (size_t)j * 3], tv, 12);
        }
    o->nmaterials = (int)mats.size();
    if (any_pr) o->rough.push_back(pr);
    for (int k = 0; k < 3; ++k) {
        o->bmin[k] = INFINITY;
        o->bmax[k] = -INFINITY;
    }
    for (size_t i = 0; i < o->pos.size(); i += 3)
        for (int k = 0; k < 3; ++k) {
            o->bmin[k] = std::min(o->bmin[k], o->pos[i + k]);
            o->bmax[k] = std::max(o->bmax[k], o->pos[i + k]);
        }
    *out = o;
    return VCT_OK;
}

int vct_obj_mesh(const vct_obj *o, vct_mesh *m) {
    if (!o || !m) return VCT_ERR_INVALID;
    m->num_tris = (int)o->albedo.size();
    m->num_objects = 1;
    m->positions = o->pos.data();
    m->normals = o->nrm.data();
    m->tangents = o->tan.data();
    m->albedo = o->albedo.data();
    m->object = o->object.data();
    m->roughness = o->rough.empty() ? nullptr : o->rough.data();
    m->vertex_tangents = o->vtan.data();
    m->texcoords = o->uv.data();
    m->texture = o->texid.data();
    return VCT_OK;
}

int vct_obj_textures(const vct_obj *o, const vct_texture **tex, int *count) {
    if (!o) return VCT_ERR_INVALID;
    if (tex) *tex = o->textures.empty() ? nullptr : o->textures.data();
    if (count) *count = (int)o->textures.size();
    return VCT_OK;
}

int vct_obj_bounds(const vct_obj *o, float bmin[3], float bmax[3], int *num_materials) {
    if (!o) return VCT_ERR_INVALID;
    if (bmin) std::memcpy(bmin, o->bmin, 12);
    if (bmax) std::memcpy(bmax, o->bmax, 12);
    if (num_materials) *num_materials = o->nmaterials;
    return VCT_OK;
}

void vct_obj_free(vct_obj *o) { delete o; }

}  // extern "C"
