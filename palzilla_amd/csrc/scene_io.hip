// Scene ingestion: Wavefront OBJ meshes as palz_triangle / palz_triangle_attrib
// arrays, with the rules of the reference's loadFrame (pk.h:1246-1362) where
// OBJ has a counterpart: positions per face corner, shading normals per face
// corner, and the material named by the mesh ("mirror" -> Mirror,
// "dielectric" -> Dielectric, any other name the caller's default, which
// loadFrame takes as Diffuse; here the name is the `usemtl` in force).
// loadFrame asserts triangles (pk.h:1320); an OBJ face of n > 3 corners is
// split into the fan (0, i, i + 1) in file order.  A corner without a `vn`
// takes the triangle's geometric normal normalize(cross(v1 - v0, v2 - v0)).
// Numbers are read as double (strtod) and rounded to float.  Host code only.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

struct palz_mesh {
    std::vector<palz_triangle> tris;
    std::vector<palz_triangle_attrib> attribs;
};

namespace palz {
namespace {

bool parse_float(const char*& s, float& out) {
    while (*s == ' ' || *s == '\t') ++s;
    char* end = nullptr;
    errno = 0;
    const double d = std::strtod(s, &end);
    if (end == s || errno == ERANGE) return false;
    s = end;
    out = (float)d;
    return true;
}

// a face corner "v", "v/vt", "v//vn" or "v/vt/vn" (1-based, negative =
// relative to the end); returns false on a malformed corner
bool parse_corner(const char*& s, long& v, long& vn, bool& has_vn) {
    char* end = nullptr;
    v = std::strtol(s, &end, 10);
    if (end == s || v == 0) return false;
    s = end;
    has_vn = false;
    if (*s == '/') {
        ++s;
        if (*s != '/') {  // texture index: read and ignore
            std::strtol(s, &end, 10);
            if (end == s) return false;
            s = end;
        }
        if (*s == '/') {
            ++s;
            vn = std::strtol(s, &end, 10);
            if (end == s || vn == 0) return false;
            s = end;
            has_vn = true;
        }
    }
    return *s == '\0' || *s == ' ' || *s == '\t' || *s == '\r' || *s == '\n';
}

bool resolve(long i, size_t n, size_t& out) {
    const long r = i > 0 ? i - 1 : (long)n + i;
    if (r < 0 || (size_t)r >= n) return false;
    out = (size_t)r;
    return true;
}

int load(const char* path, int default_material, palz_mesh& m) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(PALZ_ERR_INVALID_ARGUMENT, std::string("palz_mesh_load_obj: cannot open ") + path);
    std::vector<palz_float3> v, vn;
    int material = default_material;
    std::string line;
    char buf[4096];
    long lineno = 0;
    int rc = PALZ_OK;
    auto bad = [&](const char* what) {
        rc = fail(PALZ_ERR_INVALID_ARGUMENT,
                  std::string("palz_mesh_load_obj: ") + what + " at " + path + ":" + std::to_string(lineno));
    };
    while (rc == PALZ_OK && std::fgets(buf, sizeof(buf), f)) {
        line.assign(buf);
        while (!line.empty() && line.back() != '\n' && std::fgets(buf, sizeof(buf), f)) line.append(buf);
        ++lineno;
        const char* s = line.c_str();
        while (*s == ' ' || *s == '\t') ++s;
        auto word = [&](const char* w) {
            const size_t n = std::strlen(w);
            return std::strncmp(s, w, n) == 0 &&
                   (s[n] == ' ' || s[n] == '\t' || s[n] == '\r' || s[n] == '\n' || s[n] == '\0');
        };
        if (word("v")) {
            s += 1;
            palz_float3 p;
            if (!parse_float(s, p.x) || !parse_float(s, p.y) || !parse_float(s, p.z)) bad("bad vertex");
            v.push_back(p);
        } else if (word("vn")) {
            s += 2;
            palz_float3 p;
            if (!parse_float(s, p.x) || !parse_float(s, p.y) || !parse_float(s, p.z)) bad("bad normal");
            vn.push_back(p);
        } else if (word("usemtl")) {
            s += 6;
            while (*s == ' ' || *s == '\t') ++s;
            size_t n = std::strcspn(s, " \t\r\n");
            const std::string name(s, n);
            material = name == "mirror" ? PALZ_MATERIAL_MIRROR
                       : name == "dielectric" ? PALZ_MATERIAL_DIELECTRIC
                                              : default_material;
        } else if (word("f")) {
            s += 1;
            std::vector<size_t> cv, cn;
            std::vector<bool> hn;
            while (true) {
                while (*s == ' ' || *s == '\t') ++s;
                if (*s == '\0' || *s == '\r' || *s == '\n') break;
                long a = 0, b = 0;
                bool has = false;
                size_t ia = 0, ib = 0;
                if (!parse_corner(s, a, b, has) || !resolve(a, v.size(), ia) || (has && !resolve(b, vn.size(), ib))) {
                    bad("bad face corner");
                    break;
                }
                cv.push_back(ia);
                cn.push_back(ib);
                hn.push_back(has);
            }
            if (rc != PALZ_OK) break;
            if (cv.size() < 3) {
                bad("face with fewer than 3 corners");
                break;
            }
            for (size_t j = 1; j + 1 < cv.size(); ++j) {
                const size_t c[3] = {0, j, j + 1};
                palz_triangle t;
                palz_triangle_attrib at;
                std::memset(&at, 0, sizeof(at));
                at.material = material;
                for (int k = 0; k < 3; ++k) t.vs[k] = v[cv[c[k]]];
                const v3 g = cross(mk(t.vs[1]) - mk(t.vs[0]), mk(t.vs[2]) - mk(t.vs[0]));
                const float len = std::sqrt(dot(g, g));
                const palz_float3 ng = {g.x / len, g.y / len, g.z / len};
                for (int k = 0; k < 3; ++k) at.shading_normals[k] = hn[c[k]] ? vn[cn[c[k]]] : ng;
                m.tris.push_back(t);
                m.attribs.push_back(at);
            }
        }
        // other statements (o, g, s, vt, mtllib, l, p, comments) carry nothing used here
    }
    std::fclose(f);
    return rc;
}

}  // namespace
}  // namespace palz

extern "C" {

int palz_mesh_load_obj(const char* path, int default_material, palz_mesh** out) {
    using namespace palz;
    if (!path || !out) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_mesh_load_obj: NULL argument");
    if (default_material < PALZ_MATERIAL_DIFFUSE || default_material > PALZ_MATERIAL_DIELECTRIC)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_mesh_load_obj: unknown default material");
    palz_mesh* m = new palz_mesh();
    const int rc = load(path, default_material, *m);
    if (rc != PALZ_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return PALZ_OK;
}

int palz_mesh_triangle_count(const palz_mesh* m, uint32_t* n) {
    if (!m || !n) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_mesh_triangle_count: NULL argument");
    *n = (uint32_t)m->tris.size();
    return PALZ_OK;
}

int palz_mesh_copy(const palz_mesh* m, palz_triangle* tris, palz_triangle_attrib* attribs) {
    if (!m) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_mesh_copy: NULL mesh");
    if (tris && !m->tris.empty()) std::memcpy(tris, m->tris.data(), m->tris.size() * sizeof(palz_triangle));
    if (attribs && !m->attribs.empty())
        std::memcpy(attribs, m->attribs.data(), m->attribs.size() * sizeof(palz_triangle_attrib));
    return PALZ_OK;
}

int palz_mesh_destroy(palz_mesh* m) {
    delete m;
    return PALZ_OK;
}

}  // extern "C"
