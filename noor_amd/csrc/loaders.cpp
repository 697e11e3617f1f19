// OBJ + MTL models and float images for the payload (include/noor_loaders.h).
//
// The reference reads OBJ through glm_obj (noor/source/objLoader.cpp:28-54)
// and images through stb_image / tinyexr (noor/source/texture.cpp:49-76);
// neither is shipped with it.  This file restates glm_obj's published
// behaviour (fan triangulation, glmUnitize, glmFacetNormals,
// glmVertexNormals with a crease angle) and reads Radiance HDR, PFM and PNM.
#include "noor_loaders.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include <zlib.h>

struct noor_obj {
    std::vector<float> pos, nrm, uv;
    std::vector<uint32_t> idx, mat;
    std::vector<noor_obj_material> mats;
    std::vector<float> lights;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    float scale = 1.0f;
};

namespace {

thread_local std::string g_err;

noor_status fail(noor_status s, const std::string& m) {
    g_err = m;
    return s;
}

std::string dir_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

void copy_str(char* dst, size_t n, const std::string& s) {
    std::strncpy(dst, s.c_str(), n - 1);
    dst[n - 1] = 0;
}

noor_obj_material default_material(const std::string& name) {
    noor_obj_material m;
    std::memset(&m, 0, sizeof(m));
    copy_str(m.name, sizeof(m.name), name);
    m.type = NOOR_MAT_DIFFUSE;
    for (int c = 0; c < 3; ++c) {
        m.diffuse[c] = 0.8f;
        m.specular[c] = 0.0f;
        m.transmission[c] = 0.0f;
        m.emittance[c] = 0.0f;
    }
    m.ior = 1.0f;
    m.roughness = 1.0f / 65.0f;  // glm's default shininess
    m.transparency = 1.0f;
    m.bumpfactor = 1.0f;
    m.uvscale[0] = m.uvscale[1] = 1.0f;
    return m;
}

uint32_t illum_type(int illum) {
    switch (illum) {
        case 1: return NOOR_MAT_GLOSSY;
        case 3: return NOOR_MAT_MIRROR;
        case 5: return NOOR_MAT_EMITTER;
        case 6: return NOOR_MAT_METAL;
        case 7: return NOOR_MAT_GLASS;
        default: return NOOR_MAT_DIFFUSE;
    }
}

// "map_Kd -s 2 2 -bm 0.5 file.png": options, then the file name (rest of line)
void parse_map(std::istringstream& ls, std::string& file, float* uvscale, float* bm) {
    std::string tok;
    while (ls >> tok) {
        if (tok == "-s") {
            float a = 1, b = 1, c = 1;
            ls >> a >> b;
            // an optional third component
            const std::streampos p = ls.tellg();
            if (!(ls >> c)) {
                ls.clear();
                ls.seekg(p);
            }
            if (uvscale) {
                uvscale[0] = a;
                uvscale[1] = b;
            }
        } else if (tok == "-bm") {
            float f = 1;
            ls >> f;
            if (bm) *bm = f;
        } else if (tok == "-o" || tok == "-t") {
            float a, b;
            ls >> a >> b;
        } else if (tok.size() > 1 && tok[0] == '-') {
            float a;
            ls >> a;
        } else {
            std::string rest;
            std::getline(ls, rest);
            file = tok + rest;
            while (!file.empty() && (file.back() == ' ' || file.back() == '\r' || file.back() == '\t')) file.pop_back();
            return;
        }
    }
}

noor_status read_mtl(const std::string& path, std::vector<noor_obj_material>& mats,
                     std::map<std::string, uint32_t>& by_name) {
    std::ifstream f(path);
    if (!f) return fail(NOOR_ERR_INVALID_ARG, "cannot open material library " + path);
    std::string line;
    noor_obj_material* m = nullptr;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        std::string k;
        if (!(ls >> k) || k[0] == '#') continue;
        if (k == "newmtl") {
            std::string name;
            ls >> name;
            by_name[name] = (uint32_t)mats.size();
            mats.push_back(default_material(name));
            m = &mats.back();
            continue;
        }
        if (!m) continue;
        if (k == "Kd") ls >> m->diffuse[0] >> m->diffuse[1] >> m->diffuse[2];
        else if (k == "Ks") ls >> m->specular[0] >> m->specular[1] >> m->specular[2];
        else if (k == "Ke") ls >> m->emittance[0] >> m->emittance[1] >> m->emittance[2];
        else if (k == "Tf") ls >> m->transmission[0] >> m->transmission[1] >> m->transmission[2];
        else if (k == "Ni") ls >> m->ior;
        else if (k == "Ns") {
            float ns = 0;
            ls >> ns;
            m->roughness = 1.0f / ns;
        } else if (k == "d") ls >> m->transparency;
        else if (k == "Tr") {
            float tr = 0;
            ls >> tr;
            m->transparency = 1.0f - tr;
        } else if (k == "illum") {
            int il = 0;
            ls >> il;
            m->type = illum_type(il);
        } else {
            std::string file;
            if (k == "map_Kd") {
                parse_map(ls, file, m->uvscale, nullptr);
                copy_str(m->map_kd, NOOR_PATH_MAX, file);
            } else if (k == "map_Ks") {
                parse_map(ls, file, nullptr, nullptr);
                copy_str(m->map_ks, NOOR_PATH_MAX, file);
            } else if (k == "map_Ke") {
                parse_map(ls, file, nullptr, nullptr);
                copy_str(m->map_ke, NOOR_PATH_MAX, file);
            } else if (k == "map_d") {
                parse_map(ls, file, nullptr, nullptr);
                copy_str(m->map_d, NOOR_PATH_MAX, file);
            } else if (k == "map_bump" || k == "bump" || k == "map_Bump") {
                parse_map(ls, file, nullptr, &m->bumpfactor);
                copy_str(m->map_bump, NOOR_PATH_MAX, file);
            }
        }
    }
    return NOOR_OK;
}

// OBJ index: 1-based, negative counts back from the end; 0 = absent
bool obj_index(const std::string& s, size_t n, uint32_t& out) {
    if (s.empty()) return false;
    char* end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str()) return false;
    long i = v > 0 ? v - 1 : (long)n + v;
    if (v == 0 || i < 0 || i >= (long)n) return false;
    out = (uint32_t)i;
    return true;
}

}  // namespace

extern "C" {

const char* noor_loader_last_error(void) { return g_err.c_str(); }

static noor_status obj_load(const char* path, int unitize, float crease_deg, noor_obj** out) {
    if (!path || !out) return NOOR_ERR_INVALID_ARG;
    *out = nullptr;
    std::ifstream f(path);
    if (!f) return fail(NOOR_ERR_INVALID_ARG, std::string("cannot open ") + path);
    const std::string dir = dir_of(path);
    std::vector<float> V, N, T;                        // file arrays
    struct Corner { uint32_t v, t; bool has_t; };
    std::vector<Corner> corners;                       // 3 per triangle
    std::vector<uint32_t> tri_mat;
    struct Group { std::string name; uint32_t first_tri, count; };
    std::vector<Group> groups;
    std::vector<noor_obj_material> mats;
    std::map<std::string, uint32_t> by_name;
    uint32_t cur_mat = UINT32_MAX;
    std::string line;
    size_t lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        std::istringstream ls(line);
        std::string k;
        if (!(ls >> k) || k[0] == '#') continue;
        if (k == "v") {
            float x = 0, y = 0, z = 0;
            ls >> x >> y >> z;
            if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
                return fail(NOOR_ERR_INVALID_PAYLOAD, "non-finite vertex on line " + std::to_string(lineno));
            V.insert(V.end(), {x, y, z});
        } else if (k == "vn") {
            float x = 0, y = 0, z = 0;
            ls >> x >> y >> z;
            N.insert(N.end(), {x, y, z});
        } else if (k == "vt") {
            float u = 0, v = 0;
            ls >> u >> v;
            T.insert(T.end(), {u, v});
        } else if (k == "f") {
            std::vector<Corner> poly;
            std::string tok;
            while (ls >> tok) {
                std::string a[3];
                int part = 0;
                for (char ch : tok) {
                    if (ch == '/') {
                        if (++part > 2) break;
                    } else {
                        a[part] += ch;
                    }
                }
                Corner c{0, 0, false};
                if (!obj_index(a[0], V.size() / 3, c.v))
                    return fail(NOOR_ERR_INVALID_PAYLOAD, "bad vertex index on line " + std::to_string(lineno));
                c.has_t = obj_index(a[1], T.size() / 2, c.t);
                poly.push_back(c);
            }
            if (poly.size() < 3) continue;
            for (size_t i = 1; i + 1 < poly.size(); ++i) {
                corners.push_back(poly[0]);
                corners.push_back(poly[i]);
                corners.push_back(poly[i + 1]);
                tri_mat.push_back(cur_mat);
                if (!groups.empty()) groups.back().count++;
            }
        } else if (k == "g" || k == "o") {
            std::string name, rest;
            ls >> name;
            groups.push_back(Group{name, (uint32_t)tri_mat.size(), 0});
        } else if (k == "usemtl") {
            std::string name;
            ls >> name;
            auto it = by_name.find(name);
            cur_mat = it == by_name.end() ? UINT32_MAX : it->second;
        } else if (k == "mtllib") {
            std::string name;
            while (ls >> name) {
                const noor_status s = read_mtl(dir + name, mats, by_name);
                if (s != NOOR_OK) return s;
            }
        }
    }
    const uint32_t nt = (uint32_t)tri_mat.size();
    if (nt == 0) return fail(NOOR_ERR_INVALID_PAYLOAD, std::string("no triangles in ") + path);
    noor_obj* o = new noor_obj();
    // default material for triangles without one
    const uint32_t def = (uint32_t)mats.size();
    bool need_def = false;
    for (auto& m : tri_mat)
        if (m == UINT32_MAX) {
            m = def;
            need_def = true;
        }
    if (need_def) mats.push_back(default_material("default"));
    // glmUnitize
    const size_t nv = V.size() / 3;
    float mn[3] = {V[0], V[1], V[2]}, mx[3] = {V[0], V[1], V[2]};
    for (size_t i = 1; i < nv; ++i)
        for (int c = 0; c < 3; ++c) {
            mn[c] = std::min(mn[c], V[3 * i + c]);
            mx[c] = std::max(mx[c], V[3 * i + c]);
        }
    for (int c = 0; c < 3; ++c) {
        o->bmin[c] = mn[c];
        o->bmax[c] = mx[c];
    }
    if (unitize) {
        const float w = std::fabs(mx[0]) + std::fabs(mn[0]);
        const float h = std::fabs(mx[1]) + std::fabs(mn[1]);
        const float d = std::fabs(mx[2]) + std::fabs(mn[2]);
        const float cx = (mx[0] + mn[0]) / 2.0f, cy = (mx[1] + mn[1]) / 2.0f, cz = (mx[2] + mn[2]) / 2.0f;
        const float big = std::max(std::max(w, h), d);
        const float scale = big > 0.0f ? 2.0f / big : 1.0f;
        for (size_t i = 0; i < nv; ++i) {
            V[3 * i + 0] = (V[3 * i + 0] - cx) * scale;
            V[3 * i + 1] = (V[3 * i + 1] - cy) * scale;
            V[3 * i + 2] = (V[3 * i + 2] - cz) * scale;
        }
        o->scale = scale;
    }
    // glmFacetNormals
    std::vector<float> fn(3 * (size_t)nt);
    for (uint32_t t = 0; t < nt; ++t) {
        const float* a = &V[3 * corners[3 * t].v];
        const float* b = &V[3 * corners[3 * t + 1].v];
        const float* c = &V[3 * corners[3 * t + 2].v];
        const float u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const float v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        float n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; ++k) fn[3 * t + k] = l > 0.0f ? n[k] / l : 0.0f;
    }
    // glmVertexNormals: per vertex, the triangles that use it; the reference
    // triangle is the last one in file order (the head of glm's prepend list)
    std::vector<std::vector<uint32_t>> users(nv);
    for (uint32_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) users[corners[3 * t + k].v].push_back(t);
    const float cos_angle = std::cos(crease_deg * 3.14159265358979323846f / 180.0f);
    std::vector<float> cn(9 * (size_t)nt);   // per-corner normal
    for (size_t i = 0; i < nv; ++i) {
        const auto& U = users[i];
        if (U.empty()) continue;
        const float* ref = &fn[3 * U.back()];
        float avg[3] = {0, 0, 0};
        std::vector<bool> within(U.size());
        for (size_t j = U.size(); j-- > 0;) {  // glm walks its list from the last triangle
            const float* q = &fn[3 * U[j]];
            within[j] = q[0] * ref[0] + q[1] * ref[1] + q[2] * ref[2] > cos_angle;
            if (within[j])
                for (int k = 0; k < 3; ++k) avg[k] += q[k];
        }
        const float l = std::sqrt(avg[0] * avg[0] + avg[1] * avg[1] + avg[2] * avg[2]);
        for (int k = 0; k < 3; ++k) avg[k] = l > 0.0f ? avg[k] / l : 0.0f;
        for (size_t j = 0; j < U.size(); ++j) {
            const uint32_t t = U[j];
            for (int c = 0; c < 3; ++c) {
                if (corners[3 * t + c].v != i) continue;
                for (int k = 0; k < 3; ++k) cn[9 * (size_t)t + 3 * c + k] = within[j] ? avg[k] : fn[3 * t + k];
            }
        }
    }
    // unified vertices: one per distinct (position, normal, texcoord) of a corner
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>, uint32_t> uniq;
    for (uint32_t t = 0; t < nt; ++t)
        for (int c = 0; c < 3; ++c) {
            const Corner& k = corners[3 * t + c];
            const float* n = &cn[9 * (size_t)t + 3 * c];
            uint32_t nb[3];
            std::memcpy(nb, n, sizeof(nb));
            const uint32_t tix = k.has_t ? k.t : UINT32_MAX;
            const auto key = std::make_tuple(k.v, tix, nb[0], nb[1], nb[2]);
            auto it = uniq.find(key);
            uint32_t id;
            if (it == uniq.end()) {
                id = (uint32_t)(o->pos.size() / 3);
                uniq.emplace(key, id);
                o->pos.insert(o->pos.end(), {V[3 * k.v], V[3 * k.v + 1], V[3 * k.v + 2]});
                o->nrm.insert(o->nrm.end(), {n[0], n[1], n[2]});
                if (k.has_t) o->uv.insert(o->uv.end(), {T[2 * k.t], T[2 * k.t + 1]});
                else o->uv.insert(o->uv.end(), {0.0f, 0.0f});
            } else {
                id = it->second;
            }
            o->idx.push_back(id);
        }
    o->mat = tri_mat;
    o->mats = mats;
    // light groups (noor/source/objLoader.cpp:56-79): a quad from the first triangle
    for (const Group& g : groups) {
        if (g.name.find("light") == std::string::npos || g.count == 0) continue;
        const uint32_t t = g.first_tri;
        const float* base = &V[3 * corners[3 * t + 1].v];
        const float* v1 = &V[3 * corners[3 * t + 2].v];
        const float* v2 = &V[3 * corners[3 * t + 0].v];
        const float u[3] = {v1[0] - base[0], v1[1] - base[1], v1[2] - base[2]};
        const float v[3] = {v2[0] - base[0], v2[1] - base[1], v2[2] - base[2]};
        float n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; ++k) n[k] = l > 0.0f ? n[k] / l : 0.0f;
        const float* ke = mats[tri_mat[t]].emittance;
        o->lights.insert(o->lights.end(), {base[0], base[1], base[2], u[0], u[1], u[2], v[0], v[1], v[2], n[0], n[1],
                                           n[2], ke[0], ke[1], ke[2]});
    }
    *out = o;
    return NOOR_OK;
}

noor_status noor_obj_view_get(const noor_obj* o, noor_obj_view* v) {
    if (!o || !v) return NOOR_ERR_INVALID_ARG;
    std::memset(v, 0, sizeof(*v));
    v->num_vertices = (uint32_t)(o->pos.size() / 3);
    v->positions = o->pos.data();
    v->normals = o->nrm.data();
    v->uvs = o->uv.data();
    v->num_triangles = (uint32_t)(o->idx.size() / 3);
    v->indices = o->idx.data();
    v->material = o->mat.data();
    v->num_materials = (uint32_t)o->mats.size();
    v->materials = o->mats.data();
    v->num_lights = (uint32_t)(o->lights.size() / 15);
    v->lights = o->lights.data();
    for (int c = 0; c < 3; ++c) {
        v->bbox_min[c] = o->bmin[c];
        v->bbox_max[c] = o->bmax[c];
    }
    v->unitize_scale = o->scale;
    return NOOR_OK;
}

void noor_obj_free(noor_obj* o) { delete o; }

// ---- images -----------------------------------------------------------------------------

namespace {

bool read_all(const char* path, std::vector<unsigned char>& buf) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        return false;
    }
    buf.resize((size_t)n);
    const size_t got = n ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    return got == (size_t)n;
}

// RGBE texel -> float (Ward's Real Pixels: mantissa + 0.5 times 2^(e - 136))
void rgbe_to_float(const unsigned char* p, float* o) {
    if (p[3] == 0) {
        o[0] = o[1] = o[2] = 0.0f;
    } else {
        const float f = std::ldexp(1.0f, (int)p[3] - (128 + 8));
        o[0] = ((float)p[0] + 0.5f) * f;
        o[1] = ((float)p[1] + 0.5f) * f;
        o[2] = ((float)p[2] + 0.5f) * f;
    }
    o[3] = 1.0f;
}

// Images beyond 2^28 pixels (4 GiB of RGBA32F) are refused from the header,
// before anything is allocated.
constexpr uint64_t kMaxPixels = 1ull << 28;
bool too_large(uint64_t w, uint64_t h) { return w * h > kMaxPixels; }

noor_status load_hdr(const std::vector<unsigned char>& b, uint32_t& W, uint32_t& H, std::vector<float>& out) {
    size_t p = 0;
    auto getline_ = [&](std::string& s) {
        s.clear();
        while (p < b.size() && b[p] != '\n') s += (char)b[p++];
        if (p < b.size()) ++p;
        return !s.empty() || p < b.size();
    };
    std::string s;
    bool fmt = false;
    while (getline_(s) && !s.empty())
        if (s.rfind("FORMAT=32-bit_rle_rgbe", 0) == 0) fmt = true;
    if (!fmt) return fail(NOOR_ERR_UNSUPPORTED, "HDR: only FORMAT=32-bit_rle_rgbe");
    if (!getline_(s)) return fail(NOOR_ERR_INVALID_PAYLOAD, "HDR: no resolution line");
    char ya[4] = {0}, xa[4] = {0};
    int h = 0, w = 0;
    if (std::sscanf(s.c_str(), "%3s %d %3s %d", ya, &h, xa, &w) != 4 || std::strcmp(ya, "-Y") != 0 ||
        std::strcmp(xa, "+X") != 0 || w <= 0 || h <= 0 || w > 1 << 16 || h > 1 << 16)
        return fail(NOOR_ERR_UNSUPPORTED, "HDR: only -Y H +X W orientation");
    if (too_large(w, h)) return fail(NOOR_ERR_INVALID_PAYLOAD, "HDR: image too large");
    W = (uint32_t)w;
    H = (uint32_t)h;
    out.assign(4 * (size_t)W * H, 0.0f);
    std::vector<unsigned char> line(4 * (size_t)W);
    for (uint32_t y = 0; y < H; ++y) {
        if (p + 4 > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "HDR: truncated");
        const bool rle = W >= 8 && W < 32768 && b[p] == 2 && b[p + 1] == 2 && !(b[p + 2] & 0x80) &&
                         (((uint32_t)b[p + 2] << 8) | b[p + 3]) == W;
        if (rle) {
            p += 4;
            for (int c = 0; c < 4; ++c) {
                uint32_t x = 0;
                while (x < W) {
                    if (p >= b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "HDR: truncated run");
                    uint32_t n = b[p++];
                    if (n > 128) {
                        n -= 128;
                        if (n > W - x || p >= b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "HDR: bad run");
                        const unsigned char v = b[p++];
                        for (uint32_t k = 0; k < n; ++k) line[4 * (x++) + c] = v;
                    } else {
                        if (n == 0 || n > W - x || p + n > b.size())
                            return fail(NOOR_ERR_INVALID_PAYLOAD, "HDR: bad literal");
                        for (uint32_t k = 0; k < n; ++k) line[4 * (x++) + c] = b[p++];
                    }
                }
            }
        } else {
            if (p + 4 * (size_t)W > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "HDR: truncated");
            std::memcpy(line.data(), &b[p], 4 * (size_t)W);
            p += 4 * (size_t)W;
        }
        for (uint32_t x = 0; x < W; ++x) rgbe_to_float(&line[4 * x], &out[4 * ((size_t)y * W + x)]);
    }
    return NOOR_OK;
}

bool pnm_token(const std::vector<unsigned char>& b, size_t& p, long& v) {
    while (p < b.size()) {
        if (b[p] == '#') {
            while (p < b.size() && b[p] != '\n') ++p;
        } else if (std::isspace(b[p])) {
            ++p;
        } else {
            break;
        }
    }
    if (p >= b.size() || !std::isdigit(b[p])) return false;
    v = 0;
    while (p < b.size() && std::isdigit(b[p])) {
        v = v * 10 + (b[p++] - '0');
        if (v > 1 << 20) return false;
    }
    return true;
}

noor_status load_pnm(const std::vector<unsigned char>& b, uint32_t& W, uint32_t& H, std::vector<float>& out) {
    const int ch = b[1] == '6' ? 3 : 1;
    size_t p = 2;
    long w, h, mx;
    if (!pnm_token(b, p, w) || !pnm_token(b, p, h) || !pnm_token(b, p, mx) || w <= 0 || h <= 0 || mx <= 0 ||
        mx > 65535 || w > 1 << 16 || h > 1 << 16)
        return fail(NOOR_ERR_INVALID_PAYLOAD, "PNM: bad header");
    if (too_large(w, h)) return fail(NOOR_ERR_INVALID_PAYLOAD, "PNM: image too large");
    ++p;  // the single whitespace byte after maxval
    const int bytes = mx > 255 ? 2 : 1;
    if (p + (size_t)w * h * ch * bytes > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "PNM: truncated");
    W = (uint32_t)w;
    H = (uint32_t)h;
    out.assign(4 * (size_t)W * H, 1.0f);
    for (size_t i = 0; i < (size_t)W * H; ++i)
        for (int c = 0; c < ch; ++c) {
            long v = b[p++];
            if (bytes == 2) v = (v << 8) | b[p++];
            const float f = std::pow((float)v / (float)mx, 2.2f);
            if (ch == 1) out[4 * i] = out[4 * i + 1] = out[4 * i + 2] = f;
            else out[4 * i + c] = f;
        }
    return NOOR_OK;
}

noor_status load_pfm(const std::vector<unsigned char>& b, uint32_t& W, uint32_t& H, std::vector<float>& out) {
    const int ch = b[1] == 'F' ? 3 : 1;
    size_t p = 2;
    long w, h;
    if (!pnm_token(b, p, w) || !pnm_token(b, p, h) || w <= 0 || h <= 0 || w > 1 << 16 || h > 1 << 16)
        return fail(NOOR_ERR_INVALID_PAYLOAD, "PFM: bad header");
    if (too_large(w, h)) return fail(NOOR_ERR_INVALID_PAYLOAD, "PFM: image too large");
    while (p < b.size() && std::isspace(b[p])) ++p;
    std::string sc;
    while (p < b.size() && !std::isspace(b[p])) sc += (char)b[p++];
    ++p;
    const double scale = std::atof(sc.c_str());
    if (scale == 0.0) return fail(NOOR_ERR_INVALID_PAYLOAD, "PFM: bad scale");
    const bool little = scale < 0.0;
    if (p + (size_t)w * h * ch * 4 > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "PFM: truncated");
    W = (uint32_t)w;
    H = (uint32_t)h;
    out.assign(4 * (size_t)W * H, 1.0f);
    for (uint32_t r = 0; r < H; ++r) {
        const uint32_t y = H - 1 - r;  // stored bottom row first
        for (uint32_t x = 0; x < W; ++x)
            for (int c = 0; c < ch; ++c) {
                unsigned char q[4];
                std::memcpy(q, &b[p], 4);
                p += 4;
                if (!little) std::swap(q[0], q[3]), std::swap(q[1], q[2]);
                float f;
                std::memcpy(&f, q, 4);
                float* o = &out[4 * ((size_t)y * W + x)];
                if (ch == 1) o[0] = o[1] = o[2] = f;
                else o[c] = f;
            }
    }
    return NOOR_OK;
}

uint32_t be32(const unsigned char* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// PNG (ISO/IEC 15948): every colour type and bit depth, plain or Adam7 interlaced.
// Samples go to 8 bits first (16-bit ones keep their high byte), then to
// float as stb_image's loadf does: colour (v / 255)^2.2, alpha v / 255.
noor_status load_png(const std::vector<unsigned char>& b, uint32_t& W, uint32_t& H, std::vector<float>& out) {
    size_t p = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> idat, plte, trns;
    while (p + 8 <= b.size()) {
        const uint32_t len = be32(&b[p]);
        if (len > b.size() - p - 8 || b.size() - p - 8 - len < 4) return fail(NOOR_ERR_INVALID_PAYLOAD, "PNG: truncated chunk");
        const std::string type(reinterpret_cast<const char*>(&b[p + 4]), 4);
        const unsigned char* d = &b[p + 8];
        if (type == "IHDR") {
            if (len < 13) return fail(NOOR_ERR_INVALID_PAYLOAD, "PNG: bad IHDR");
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (type == "PLTE") {
            plte.assign(d, d + len);
        } else if (type == "tRNS") {
            trns.assign(d, d + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            break;
        }
        p += 12 + (size_t)len;
    }
    const int chans = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    const bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                          (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                          ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!chans || !depth_ok || w == 0 || h == 0 || w > 1 << 16 || h > 1 << 16)
        return fail(NOOR_ERR_INVALID_PAYLOAD, "PNG: bad header");
    if (too_large(w, h)) return fail(NOOR_ERR_INVALID_PAYLOAD, "PNG: image too large");
    if (interlace > 1) return fail(NOOR_ERR_INVALID_PAYLOAD, "PNG: bad interlace method");
    if (ctype == 3 && (plte.empty() || plte.size() % 3)) return fail(NOOR_ERR_INVALID_PAYLOAD, "PNG: no palette");
    const size_t bits = (size_t)chans * depth, bpp = std::max<size_t>(1, bits / 8);
    // passes: (x0, y0, dx, dy); one pass covering everything when not interlaced (Adam7 otherwise)
    static const int adam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                    {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    static const int flat[1][4] = {{0, 0, 1, 1}};
    const int npass = interlace ? 7 : 1;
    const int (*passes)[4] = interlace ? adam7 : flat;
    size_t total = 0;
    for (int k = 0; k < npass; ++k) {
        const uint32_t pw = w > (uint32_t)passes[k][0] ? (w - passes[k][0] + passes[k][2] - 1) / passes[k][2] : 0;
        const uint32_t ph = h > (uint32_t)passes[k][1] ? (h - passes[k][1] + passes[k][3] - 1) / passes[k][3] : 0;
        if (pw && ph) total += (((size_t)pw * bits + 7) / 8 + 1) * ph;
    }
    std::vector<unsigned char> raw(total);
    uLongf n = (uLongf)raw.size();
    if (uncompress(raw.data(), &n, idat.data(), (uLong)idat.size()) != Z_OK || n != raw.size())
        return fail(NOOR_ERR_INVALID_PAYLOAD, "PNG: bad image data");
    // 8-bit samples (palette indices for colour type 3) of every texel
    std::vector<int> smp((size_t)w * h * chans);
    size_t off = 0;
    for (int k = 0; k < npass; ++k) {
        const uint32_t x0 = passes[k][0], y0 = passes[k][1], dx = passes[k][2], dy = passes[k][3];
        const uint32_t pw = w > x0 ? (w - x0 + dx - 1) / dx : 0, ph = h > y0 ? (h - y0 + dy - 1) / dy : 0;
        if (!pw || !ph) continue;
        const size_t stride = ((size_t)pw * bits + 7) / 8;
        // unfilter (filters 0-4: none, sub, up, average, Paeth)
        std::vector<unsigned char> cur(stride), prev(stride, 0);
        for (uint32_t y = 0; y < ph; ++y) {
            const unsigned char f = raw[off];
            const unsigned char* src = &raw[off + 1];
            off += stride + 1;
            for (size_t i = 0; i < stride; ++i) {
                const int a = i >= bpp ? cur[i - bpp] : 0, up = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
                int v = src[i];
                switch (f) {
                    case 0: break;
                    case 1: v += a; break;
                    case 2: v += up; break;
                    case 3: v += (a + up) >> 1; break;
                    case 4: {
                        const int pp = a + up - c, pa = std::abs(pp - a), pb = std::abs(pp - up), pc = std::abs(pp - c);
                        v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? up : c);
                        break;
                    }
                    default: return fail(NOOR_ERR_INVALID_PAYLOAD, "PNG: bad filter");
                }
                cur[i] = (unsigned char)v;
            }
            for (uint32_t x = 0; x < pw; ++x)
                for (int ch = 0; ch < chans; ++ch) {
                    int v;
                    if (depth == 16) {
                        v = cur[2 * ((size_t)x * chans + ch)];  // the high byte
                    } else if (depth == 8) {
                        v = cur[(size_t)x * chans + ch];
                    } else {
                        const size_t bit = (size_t)x * depth;
                        v = (cur[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1 << depth) - 1);
                        if (ctype != 3) v *= 255 / ((1 << depth) - 1);
                    }
                    smp[((size_t)(y0 + y * dy) * w + (x0 + x * dx)) * chans + ch] = v;
                }
            std::swap(cur, prev);
        }
    }
    W = w;
    H = h;
    out.assign(4 * (size_t)w * h, 1.0f);
    auto sample = [&](uint32_t y, uint32_t x, int c) -> int { return smp[((size_t)y * w + x) * chans + c]; };
    const float g = 2.2f;
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            float* o = &out[4 * ((size_t)y * w + x)];
            if (ctype == 3) {
                const int i = sample(y, x, 0);
                if ((size_t)i * 3 + 2 >= plte.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "PNG: palette index");
                for (int c = 0; c < 3; ++c) o[c] = std::pow(plte[3 * i + c] / 255.0f, g);
                o[3] = (size_t)i < trns.size() ? trns[i] / 255.0f : 1.0f;
            } else if (chans <= 2) {
                const float v = std::pow(sample(y, x, 0) / 255.0f, g);
                o[0] = o[1] = o[2] = v;
                if (chans == 2) o[3] = sample(y, x, 1) / 255.0f;
            } else {
                for (int c = 0; c < 3; ++c) o[c] = std::pow(sample(y, x, c) / 255.0f, g);
                if (chans == 4) o[3] = sample(y, x, 3) / 255.0f;
            }
        }
    }
    return NOOR_OK;
}

// Truevision TGA: uncompressed or run-length, true-colour (24/32 bit), grey
// (8 bit) or colour-mapped (8-bit indices into a 24/32-bit map); converted
// as stb_image's loadf does (colour (v / 255)^2.2, alpha v / 255).
noor_status load_tga(const std::vector<unsigned char>& b, uint32_t& W, uint32_t& H, std::vector<float>& out) {
    if (b.size() < 18) return fail(NOOR_ERR_INVALID_PAYLOAD, "TGA: truncated header");
    const int idlen = b[0], cmtype = b[1], itype = b[2];
    const int cmfirst = b[3] | b[4] << 8, cmlen = b[5] | b[6] << 8, cmbits = b[7];
    const int w = b[12] | b[13] << 8, h = b[14] | b[15] << 8, bits = b[16], desc = b[17];
    const bool rle = itype >= 9;
    const int base = rle ? itype - 8 : itype;
    if (!(base == 1 || base == 2 || base == 3) || w == 0 || h == 0)
        return fail(NOOR_ERR_UNSUPPORTED, "TGA: only colour-mapped, true-colour and grey images");
    if (too_large(w, h)) return fail(NOOR_ERR_INVALID_PAYLOAD, "TGA: image too large");
    const bool mapped = base == 1;
    if ((mapped && (cmtype != 1 || bits != 8 || (cmbits != 24 && cmbits != 32))) ||
        (base == 2 && bits != 24 && bits != 32) || (base == 3 && bits != 8))
        return fail(NOOR_ERR_UNSUPPORTED, "TGA: unsupported pixel depth");
    size_t p = 18 + (size_t)idlen;
    std::vector<unsigned char> cmap;
    if (cmtype == 1) {
        const size_t n = (size_t)cmlen * ((cmbits + 7) / 8);
        if (p + n > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "TGA: truncated colour map");
        cmap.assign(b.begin() + p, b.begin() + p + n);
        p += n;
    }
    const int bpp = bits / 8, cbpp = cmbits / 8;
    std::vector<unsigned char> px((size_t)w * h * bpp);
    if (!rle) {
        if (p + px.size() > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "TGA: truncated pixels");
        std::memcpy(px.data(), &b[p], px.size());
    } else {
        size_t o = 0;
        while (o < px.size()) {
            if (p >= b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "TGA: truncated run");
            const int c = b[p++];
            const size_t n = (size_t)(c & 127) + 1;
            if (o + n * bpp > px.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "TGA: run overflows the image");
            if (c & 128) {
                if (p + bpp > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "TGA: truncated run");
                for (size_t k = 0; k < n; ++k) std::memcpy(&px[o + k * bpp], &b[p], bpp);
                p += bpp;
            } else {
                if (p + n * bpp > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "TGA: truncated run");
                std::memcpy(&px[o], &b[p], n * bpp);
                p += n * bpp;
            }
            o += n * bpp;
        }
    }
    W = (uint32_t)w;
    H = (uint32_t)h;
    out.assign(4 * (size_t)w * h, 1.0f);
    const bool top = (desc & 0x20) != 0, right = (desc & 0x10) != 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const unsigned char* s = &px[((size_t)y * w + x) * bpp];
            const int oy = top ? y : h - 1 - y, ox = right ? w - 1 - x : x;
            float* o = &out[4 * ((size_t)oy * w + ox)];
            int r, g, bl, a = 255, ch;
            if (mapped) {
                const int i = s[0] - cmfirst;
                if (i < 0 || (size_t)(i + 1) * cbpp > cmap.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "TGA: colour index");
                const unsigned char* c = &cmap[(size_t)i * cbpp];
                bl = c[0], g = c[1], r = c[2];
                if (cbpp == 4) a = c[3];
                ch = cbpp;
            } else if (base == 3) {
                r = g = bl = s[0];
                ch = 1;
            } else {
                bl = s[0], g = s[1], r = s[2];
                if (bpp == 4) a = s[3];
                ch = bpp;
            }
            o[0] = std::pow(r / 255.0f, 2.2f);
            o[1] = std::pow(g / 255.0f, 2.2f);
            o[2] = std::pow(bl / 255.0f, 2.2f);
            if (ch == 4) o[3] = a / 255.0f;
        }
    return NOOR_OK;
}

// Baseline JPEG (ITU T.81 sequential DCT, Huffman, 8-bit samples): one
// (grey) or three (YCbCr, JFIF) components, sampling factors 1 or 2, restart
// intervals.  Chroma is upsampled by replication (stb_image interpolates
// 2x chroma; the two differ by at most the chroma step of one texel).
// Converted as stb_image's loadf does: (v / 255)^2.2.
struct JpegHuff {
    uint8_t bits[17] = {0}, vals[256] = {0};
    int mincode[17] = {0}, maxcode[18] = {0}, valptr[17] = {0};
    bool ok = false;
};

struct JpegBits {
    const std::vector<unsigned char>& b;
    size_t p;
    uint32_t acc = 0;
    int n = 0;
    bool marker = false;
    int starved = 0;  // bytes fed as zeros since the last restart (the scan ran short)
    JpegBits(const std::vector<unsigned char>& buf, size_t pos) : b(buf), p(pos) {}
    int bit() {
        if (n == 0) {
            unsigned char c = 0;
            if (marker || p >= b.size()) ++starved;
            if (!marker && p < b.size()) {
                c = b[p++];
                if (c == 0xFF) {
                    const unsigned char d = p < b.size() ? b[p] : 0;
                    if (d == 0) ++p;           // stuffed zero
                    else { marker = true; c = 0; --p; ++starved; }  // a marker: feed zeros
                }
            }
            acc = c;
            n = 8;
        }
        --n;
        return (acc >> n) & 1;
    }
    int bits(int k) {
        int v = 0;
        for (int i = 0; i < k; ++i) v = (v << 1) | bit();
        return v;
    }
    void reset() {  // at a restart marker
        n = 0;
        marker = false;
        starved = 0;
        while (p + 1 < b.size() && !(b[p] == 0xFF && b[p + 1] >= 0xD0 && b[p + 1] <= 0xD7)) ++p;
        if (p + 1 < b.size()) p += 2;
    }
};

static int jpeg_decode_huff(JpegBits& bs, const JpegHuff& h) {
    int code = 0;
    for (int l = 1; l <= 16; ++l) {
        code = (code << 1) | bs.bit();
        if (h.maxcode[l] >= 0 && code <= h.maxcode[l] && code >= h.mincode[l])
            return h.vals[h.valptr[l] + code - h.mincode[l]];
    }
    return -1;
}

static int jpeg_extend(int v, int t) { return t && v < (1 << (t - 1)) ? v - (1 << t) + 1 : v; }

noor_status load_jpeg(const std::vector<unsigned char>& b, uint32_t& W, uint32_t& H, std::vector<float>& out) {
    static const int zz[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                               12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                               35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                               58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
    uint16_t qt[4][64] = {{0}};
    JpegHuff hd[4], ha[4];
    double idct[8][8];  // C(u) cos((2x + 1) u pi / 16)
    for (int x = 0; x < 8; ++x)
        for (int u = 0; u < 8; ++u)
            idct[x][u] = (u ? 1.0 : 0.70710678118654752) * std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16);
    struct Comp { int id, h, v, tq, td = 0, ta = 0; std::vector<uint8_t> px; int bw = 0, bh = 0; };
    std::vector<Comp> comps;
    int w = 0, h = 0, restart = 0;
    size_t p = 2;
    auto u16 = [&](size_t at) { return (int)b[at] << 8 | b[at + 1]; };
    while (p + 4 <= b.size()) {
        if (b[p] != 0xFF) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: expected a marker");
        const int m = b[p + 1];
        if (m == 0xD8 || (m >= 0xD0 && m <= 0xD7) || m == 0x01 || m == 0xFF) { p += (m == 0xFF) ? 1 : 2; continue; }
        if (m == 0xD9) break;
        const int len = u16(p + 2);
        if (len < 2 || p + 2 + (size_t)len > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: bad segment");
        const size_t q0 = p + 4, q1 = p + 2 + len;
        if (m == 0xC2 || m == 0xC6 || m == 0xCA || m == 0xCE || (m >= 0xC3 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC))
            return fail(NOOR_ERR_UNSUPPORTED, "JPEG: only baseline (sequential Huffman) images");
        if (m == 0xDB) {  // quantisation tables
            size_t q = q0;
            while (q < q1) {
                const int pq = b[q] >> 4, tq = b[q] & 3;
                ++q;
                for (int i = 0; i < 64; ++i) {
                    if (q + (pq ? 2 : 1) > q1) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: bad DQT");
                    qt[tq][zz[i]] = pq ? (uint16_t)u16(q) : b[q];
                    q += pq ? 2 : 1;
                }
            }
        } else if (m == 0xC4) {  // Huffman tables
            size_t q = q0;
            while (q < q1) {
                const int tc = b[q] >> 4, th = b[q] & 3;
                JpegHuff& t = tc ? ha[th] : hd[th];
                ++q;
                int total = 0;
                for (int l = 1; l <= 16; ++l) {
                    if (q >= q1) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: bad DHT");
                    t.bits[l] = b[q++];
                    total += t.bits[l];
                }
                if (total > 256 || q + total > q1) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: bad DHT");
                for (int i = 0; i < total; ++i) t.vals[i] = b[q++];
                int code = 0, k = 0;
                for (int l = 1; l <= 16; ++l) {
                    t.valptr[l] = k;
                    t.mincode[l] = code;
                    code += t.bits[l];
                    k += t.bits[l];
                    t.maxcode[l] = t.bits[l] ? code - 1 : -1;
                    code <<= 1;
                }
                t.ok = true;
            }
        } else if (m == 0xDD) {
            if (q0 + 2 > q1) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: bad DRI");
            restart = u16(q0);
        } else if (m == 0xC0 || m == 0xC1) {
            if (q0 + 6 > q1 || !comps.empty()) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: bad frame header");
            if (b[q0] != 8) return fail(NOOR_ERR_UNSUPPORTED, "JPEG: only 8-bit samples");
            h = u16(q0 + 1);
            w = u16(q0 + 3);
            const int nc = b[q0 + 5];
            if ((nc != 1 && nc != 3) || q0 + 6 + 3 * (size_t)nc > q1 || w == 0 || h == 0)
                return fail(NOOR_ERR_UNSUPPORTED, "JPEG: only grey or three-component images");
            if (too_large(w, h)) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: image too large");
            for (int i = 0; i < nc; ++i) {
                Comp c;
                c.id = b[q0 + 6 + 3 * i];
                c.h = b[q0 + 7 + 3 * i] >> 4;
                c.v = b[q0 + 7 + 3 * i] & 15;
                c.tq = b[q0 + 8 + 3 * i] & 3;
                if (c.h < 1 || c.h > 2 || c.v < 1 || c.v > 2) return fail(NOOR_ERR_UNSUPPORTED, "JPEG: sampling factors");
                comps.push_back(c);
            }
        } else if (m == 0xDA) {  // start of scan: the entropy-coded data follows the header
            if (comps.empty()) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: scan before frame");
            if (q0 + 1 > q1 || q0 + 1 + 2 * (size_t)b[q0] + 3 > q1) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: bad scan header");
            const int ns = b[q0];
            if (ns != (int)comps.size()) return fail(NOOR_ERR_UNSUPPORTED, "JPEG: only interleaved single scans");
            for (int i = 0; i < ns; ++i) {
                const int id = b[q0 + 1 + 2 * i], t = b[q0 + 2 + 2 * i];
                for (Comp& c : comps)
                    if (c.id == id) {
                        c.td = t >> 4 & 3;
                        c.ta = t & 3;
                    }
            }
            // a scan of one component is not interleaved: one 8x8 block per MCU in
            // raster order whatever its sampling factors say (T.81 A.2.2)
            if (comps.size() == 1) comps[0].h = comps[0].v = 1;
            int hmax = 1, vmax = 1;
            for (const Comp& c : comps) {
                hmax = std::max(hmax, c.h);
                vmax = std::max(vmax, c.v);
            }
            const int mcux = (w + 8 * hmax - 1) / (8 * hmax), mcuy = (h + 8 * vmax - 1) / (8 * vmax);
            for (Comp& c : comps) {
                if (!hd[c.td].ok || !ha[c.ta].ok) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: missing Huffman table");
                c.bw = mcux * c.h * 8;
                c.bh = mcuy * c.v * 8;
                c.px.assign((size_t)c.bw * c.bh, 0);
            }
            JpegBits bs(b, q1);
            std::vector<int> pred(comps.size(), 0);
            int mcus = 0;
            double blk[64], tmp[64];
            for (int my = 0; my < mcuy; ++my)
                for (int mx = 0; mx < mcux; ++mx) {
                    if (restart && mcus && mcus % restart == 0) {
                        bs.reset();
                        std::fill(pred.begin(), pred.end(), 0);
                    }
                    ++mcus;
                    for (size_t ci = 0; ci < comps.size(); ++ci) {
                        Comp& c = comps[ci];
                        for (int by = 0; by < c.v; ++by)
                            for (int bx = 0; bx < c.h; ++bx) {
                                int coef[64] = {0};
                                const int t = jpeg_decode_huff(bs, hd[c.td]);
                                if (t < 0 || t > 11) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: bad DC code");
                                pred[ci] += jpeg_extend(bs.bits(t), t);
                                coef[0] = pred[ci] * qt[c.tq][0];
                                for (int k = 1; k < 64;) {
                                    const int rs = jpeg_decode_huff(bs, ha[c.ta]);
                                    if (rs < 0) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: bad AC code");
                                    const int r = rs >> 4, sz = rs & 15;
                                    if (sz == 0) {
                                        if (r != 15) break;  // end of block
                                        k += 16;
                                        continue;
                                    }
                                    k += r;
                                    if (k > 63) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: AC index");
                                    coef[zz[k]] = jpeg_extend(bs.bits(sz), sz) * qt[c.tq][zz[k]];
                                    ++k;
                                }
                                if (bs.starved > 1) return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: truncated scan");
                                // separable inverse DCT in binary64: f = 1/4 sum C(u) C(v) F cos cos
                                for (int y = 0; y < 8; ++y)
                                    for (int x = 0; x < 8; ++x) {
                                        double acc = 0;
                                        for (int u = 0; u < 8; ++u) acc += idct[x][u] * coef[y * 8 + u];
                                        tmp[y * 8 + x] = acc / 2;
                                    }
                                for (int x = 0; x < 8; ++x)
                                    for (int y = 0; y < 8; ++y) {
                                        double acc = 0;
                                        for (int v = 0; v < 8; ++v) acc += idct[y][v] * tmp[v * 8 + x];
                                        blk[y * 8 + x] = acc / 2;
                                    }
                                const int ox = (mx * c.h + bx) * 8, oy = (my * c.v + by) * 8;
                                for (int y = 0; y < 8; ++y)
                                    for (int x = 0; x < 8; ++x) {
                                        const long v = std::lround(blk[y * 8 + x] + 128.0);
                                        c.px[(size_t)(oy + y) * c.bw + ox + x] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
                                    }
                            }
                    }
                }
            W = (uint32_t)w;
            H = (uint32_t)h;
            out.assign(4 * (size_t)W * H, 1.0f);
            for (uint32_t y = 0; y < H; ++y)
                for (uint32_t x = 0; x < W; ++x) {
                    auto at = [&](const Comp& c) {
                        const size_t sx = (size_t)x * c.h / hmax, sy = (size_t)y * c.v / vmax;
                        return (double)c.px[sy * c.bw + sx];
                    };
                    float* o = &out[4 * ((size_t)y * W + x)];
                    if (comps.size() == 1) {
                        o[0] = o[1] = o[2] = std::pow((float)(at(comps[0]) / 255.0), 2.2f);
                    } else {  // JFIF YCbCr -> RGB
                        const double Y = at(comps[0]), cb = at(comps[1]) - 128.0, cr = at(comps[2]) - 128.0;
                        const double rgb[3] = {Y + 1.402 * cr, Y - 0.344136 * cb - 0.714136 * cr, Y + 1.772 * cb};
                        for (int k = 0; k < 3; ++k) {
                            const double v = std::round(rgb[k]);
                            o[k] = std::pow((float)((v < 0 ? 0 : (v > 255 ? 255 : v)) / 255.0), 2.2f);
                        }
                    }
                }
            return NOOR_OK;
        }
        p = q1;
    }
    return fail(NOOR_ERR_INVALID_PAYLOAD, "JPEG: no image data");
}

float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 31, m = h & 1023;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = sign;
        } else {  // subnormal: m * 2^-24
            const float f = std::ldexp((float)m, -24);
            std::memcpy(&bits, &f, 4);
            bits |= sign;
        }
    } else if (e == 31) {
        bits = sign | 0x7F800000u | (m << 13);
    } else {
        bits = sign | ((e + 112) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// OpenEXR (the reference reads .exr through tinyexr, noor/source/texture.cpp:51-56):
// single-part scanline images, channels R G B (A, or Y for grey) in HALF or
// FLOAT, compression NONE, RLE, ZIPS or ZIP.
noor_status load_exr(const std::vector<unsigned char>& b, uint32_t& W, uint32_t& H, std::vector<float>& out) {
    // version flags: tiled (bit 9), deep data (bit 11), multi-part (bit 12)
    if (b.size() < 8 || (b[5] & 0x1A)) return fail(NOOR_ERR_UNSUPPORTED, "EXR: only single-part scanline files");
    size_t p = 8;
    struct Ch { std::string name; int type; };
    std::vector<Ch> chans;
    int comp = -1;
    int32_t win[4] = {0, 0, -1, -1};
    auto rd32 = [&](size_t at) { int32_t v; std::memcpy(&v, &b[at], 4); return v; };
    for (;;) {
        if (p >= b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: truncated header");
        if (b[p] == 0) { ++p; break; }
        std::string name, type;
        while (p < b.size() && b[p]) name += (char)b[p++];
        ++p;
        while (p < b.size() && b[p]) type += (char)b[p++];
        ++p;
        if (p + 4 > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: truncated header");
        const int32_t size = rd32(p);
        p += 4;
        if (size < 0 || p + (size_t)size > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: bad attribute");
        const size_t end = p + (size_t)size;
        if (name == "channels") {
            size_t q = p;
            while (q < end && b[q]) {
                Ch c;
                while (q < end && b[q]) c.name += (char)b[q++];
                ++q;
                if (q + 16 > end) return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: bad channel list");
                c.type = rd32(q);
                if (rd32(q + 8) != 1 || rd32(q + 12) != 1) return fail(NOOR_ERR_UNSUPPORTED, "EXR: subsampled channels");
                q += 16;
                chans.push_back(c);
            }
        } else if (name == "compression" && size >= 1) {
            comp = b[p];
        } else if (name == "dataWindow" && size >= 16) {
            for (int k = 0; k < 4; ++k) win[k] = rd32(p + 4 * k);
        }
        p = end;
    }
    const int64_t w = (int64_t)win[2] - win[0] + 1, h = (int64_t)win[3] - win[1] + 1;
    if (chans.empty() || w <= 0 || h <= 0 || w > 1 << 16 || h > 1 << 16)
        return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: bad header");
    if (too_large(w, h)) return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: image too large");
    if (comp < 0 || comp > 3) return fail(NOOR_ERR_UNSUPPORTED, "EXR: compression other than NONE, RLE, ZIPS, ZIP");
    size_t line_bytes = 0;
    for (const Ch& c : chans) {
        if (c.type != 1 && c.type != 2) return fail(NOOR_ERR_UNSUPPORTED, "EXR: only HALF and FLOAT channels");
        line_bytes += (size_t)w * (c.type == 1 ? 2 : 4);
    }
    const int lines = comp == 3 ? 16 : 1;
    const size_t nchunks = (size_t)((h + lines - 1) / lines);
    if (p + 8 * nchunks > b.size()) return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: truncated offsets");
    W = (uint32_t)w;
    H = (uint32_t)h;
    out.assign(4 * (size_t)W * H, 0.0f);
    for (size_t i = 0; i < (size_t)W * H; ++i) out[4 * i + 3] = 1.0f;
    std::vector<unsigned char> buf, tmp;
    for (size_t k = 0; k < nchunks; ++k) {
        uint64_t off;
        std::memcpy(&off, &b[p + 8 * k], 8);
        // compared as sizes so that an offset near 2^64 cannot wrap
        if (off > b.size() || b.size() - off < 8) return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: bad chunk offset");
        const int32_t y0 = rd32((size_t)off) - win[1], size = rd32((size_t)off + 4);
        if (y0 < 0 || y0 >= h || size < 0 || (uint64_t)size > b.size() - off - 8)
            return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: bad chunk");
        const int ny = (int)std::min<int64_t>(lines, h - y0);
        const size_t want = line_bytes * ny;
        const unsigned char* data = &b[(size_t)off + 8];
        if (comp == 0 || (size_t)size == want) {  // stored uncompressed (also when compression did not pay)
            if ((size_t)size != want) return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: bad chunk size");
            buf.assign(data, data + want);
        } else {
            tmp.assign(want, 0);
            if (comp == 1) {  // RLE: count < 0 -> -count literal bytes, else count + 1 copies
                size_t q = 0, o = 0;
                while (q < (size_t)size && o < want) {
                    const int c = (signed char)data[q++];
                    if (c < 0) {
                        if (q + (size_t)(-c) > (size_t)size || o + (size_t)(-c) > want)
                            return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: bad RLE");
                        std::memcpy(&tmp[o], &data[q], (size_t)(-c));
                        q += (size_t)(-c);
                        o += (size_t)(-c);
                    } else {
                        if (q >= (size_t)size || o + (size_t)c + 1 > want) return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: bad RLE");
                        std::memset(&tmp[o], data[q++], (size_t)c + 1);
                        o += (size_t)c + 1;
                    }
                }
                if (o != want) return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: short RLE chunk");
            } else {
                uLongf n = (uLongf)want;
                if (uncompress(tmp.data(), &n, data, (uLong)size) != Z_OK || n != want)
                    return fail(NOOR_ERR_INVALID_PAYLOAD, "EXR: bad ZIP chunk");
            }
            // undo the predictor, then the split into even and odd bytes
            for (size_t i = 1; i < want; ++i) tmp[i] = (unsigned char)(tmp[i - 1] + tmp[i] - 128);
            buf.resize(want);
            const size_t half = (want + 1) / 2;
            for (size_t i = 0; i < want; ++i) buf[i] = (i & 1) ? tmp[half + i / 2] : tmp[i / 2];
        }
        size_t q = 0;
        for (int yy = 0; yy < ny; ++yy) {
            float* row = &out[4 * (size_t)(y0 + yy) * W];
            for (const Ch& c : chans) {
                const int slot = c.name == "R" ? 0 : c.name == "G" ? 1 : c.name == "B" ? 2 : c.name == "A" ? 3
                                                                                               : c.name == "Y" ? 4 : -1;
                for (uint32_t x = 0; x < W; ++x) {
                    float v;
                    if (c.type == 1) {
                        uint16_t hv;
                        std::memcpy(&hv, &buf[q], 2);
                        q += 2;
                        v = half_to_float(hv);
                    } else {
                        std::memcpy(&v, &buf[q], 4);
                        q += 4;
                    }
                    if (slot == 4) row[4 * x] = row[4 * x + 1] = row[4 * x + 2] = v;
                    else if (slot >= 0) row[4 * x + slot] = v;
                }
            }
        }
    }
    return NOOR_OK;
}

}  // namespace

static noor_status image_load(const char* path, uint32_t* width, uint32_t* height, float** rgba) {
    if (!path || !width || !height || !rgba) return NOOR_ERR_INVALID_ARG;
    *rgba = nullptr;
    std::vector<unsigned char> b;
    if (!read_all(path, b)) return fail(NOOR_ERR_INVALID_ARG, std::string("cannot read ") + path);
    std::vector<float> px;
    uint32_t W = 0, H = 0;
    noor_status s;
    if (b.size() >= 2 && b[0] == '#' && b[1] == '?') s = load_hdr(b, W, H, px);
    else if (b.size() >= 2 && b[0] == 'P' && (b[1] == '5' || b[1] == '6')) s = load_pnm(b, W, H, px);
    else if (b.size() >= 2 && b[0] == 'P' && (b[1] == 'F' || b[1] == 'f')) s = load_pfm(b, W, H, px);
    else if (b.size() >= 8 && std::memcmp(b.data(), "\x89PNG\r\n\x1a\n", 8) == 0) s = load_png(b, W, H, px);
    else if (b.size() >= 8 && b[0] == 0x76 && b[1] == 0x2f && b[2] == 0x31 && b[3] == 0x01) s = load_exr(b, W, H, px);
    else if (b.size() >= 4 && b[0] == 0xFF && b[1] == 0xD8 && b[2] == 0xFF) s = load_jpeg(b, W, H, px);
    else if (b.size() >= 18 && ((b[2] >= 1 && b[2] <= 3) || (b[2] >= 9 && b[2] <= 11)) && b[1] <= 1 &&
             std::string(path).size() >= 4 &&
             (std::string(path).compare(std::string(path).size() - 4, 4, ".tga") == 0 ||
              std::string(path).compare(std::string(path).size() - 4, 4, ".TGA") == 0))
        s = load_tga(b, W, H, px);  // TGA has no signature: by extension
    else s = fail(NOOR_ERR_UNSUPPORTED, std::string("unsupported image format: ") + path);
    if (s != NOOR_OK) return s;
    float* o = new float[px.size()];
    std::memcpy(o, px.data(), px.size() * sizeof(float));
    *width = W;
    *height = H;
    *rgba = o;
    return NOOR_OK;
}

// The C entry points: no C++ exception crosses them (an allocation that
// fails becomes NOOR_ERR_OUT_OF_MEMORY).
noor_status noor_obj_load(const char* path, int unitize, float crease_deg, noor_obj** out) {
    try {
        return obj_load(path, unitize, crease_deg, out);
    } catch (const std::bad_alloc&) {
        return fail(NOOR_ERR_OUT_OF_MEMORY, "OBJ: out of memory");
    } catch (...) {
        return fail(NOOR_ERR_INVALID_PAYLOAD, "OBJ: unreadable file");
    }
}

noor_status noor_image_load(const char* path, uint32_t* width, uint32_t* height, float** rgba) {
    try {
        return image_load(path, width, height, rgba);
    } catch (const std::bad_alloc&) {
        return fail(NOOR_ERR_OUT_OF_MEMORY, "image: out of memory");
    } catch (...) {
        return fail(NOOR_ERR_INVALID_PAYLOAD, "image: unreadable file");
    }
}

void noor_image_free(float* rgba) { delete[] rgba; }

}  // extern "C"
