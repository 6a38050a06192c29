// Model-file ingestion: the reference's mesh queue (inputs.rs:93-129) built
// from OBJ and STL files, host side.
//
// OBJ follows tobj 3.2.2 (Cargo.lock:526-532) as inputs.rs:108 calls it, with
// GPU_LOAD_OPTIONS (single index, triangulate, points and lines ignored), then
// to_simple_mesh_with_materials (geometry.rs:83-138):
//   * one model per run of faces; `o` / `g` end the current model, and
//     `usemtl` ends it when the material changes (a model keeps the material
//     current when it ends);
//   * faces fan-triangulate from their first corner: (a,b,c), (a,c,d), ...;
//     face indices are 1-based, negative ones count back from the last `v`;
//   * `v x y z r g b` carries a vertex colour;
//   * the mtllib files (paths relative to the OBJ) give each `newmtl` its
//     `Kd` (default 0,0,0);
//   * colour = (Kd * 255) as u8 of the model's material, or, when the file
//     has vertex colours, (c * 255) as u8 of the triangle's first corner; with
//     no materials at all every triangle keeps (1,1,1) (geometry.rs:90-97);
//   * bounding box from the origin, widened by each triangle's aabb
//     (geometry.rs:85-88, 129-135).
// STL follows stl_io 0.4.2 (Cargo.lock:502-506) read_stl (inputs.rs:117) and
// to_simple_mesh (geometry.rs:146-181): ASCII when the first line starts with
// "solid ", else binary (80-byte header, u32 count, 50-byte facets); colour
// (255,255,0); bounding box from (f32::MAX, f32::MIN).
//
// A file the reference would panic on (no materials library where one is
// named, a model without a material while materials exist, bad indices)
// returns an error instead.
#include "sloth_models.h"

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

namespace sloth {

namespace {

// Rust's `f as u8`: saturating, NaN -> 0.
uint8_t as_u8(float f)
{
    if (!(f > 0.0f)) return 0;
    if (f >= 255.0f) return 255;
    return (uint8_t)f;
}

// Rust's f32::from_str on one token (decimal, inf, infinity, nan; no hex).
bool parse_f32(const std::string& t, float& v)
{
    if (t.empty() || t.find_first_of("xX") != std::string::npos) return false;
    char* end = nullptr;
    errno = 0;
    v = strtof(t.c_str(), &end);
    return end == t.c_str() + t.size();
}

bool parse_int(const std::string& t, long long& v)
{
    if (t.empty()) return false;
    char* end = nullptr;
    errno = 0;
    v = strtoll(t.c_str(), &end, 10);
    return end == t.c_str() + t.size() && errno == 0;
}

std::vector<std::string> words(const std::string& line)
{
    std::vector<std::string> w;
    std::istringstream is(line);
    std::string t;
    while (is >> t) w.push_back(t);
    return w;
}

std::string rest_of(const std::string& line, const std::string& key)
{
    size_t p = line.find(key);
    std::string r = line.substr(p + key.size());
    size_t a = r.find_first_not_of(" \t\r\n"), b = r.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : r.substr(a, b - a + 1);
}

bool read_file(const std::string& path, std::string& data)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    data = ss.str();
    return true;
}

std::vector<std::string> split_lines(const std::string& data)
{
    std::vector<std::string> lines;
    size_t s = 0;
    while (s <= data.size()) {
        size_t e = data.find('\n', s);
        if (e == std::string::npos) e = data.size();
        std::string l = data.substr(s, e - s);
        if (!l.empty() && l.back() == '\r') l.pop_back();
        if (e < data.size() || !l.empty()) lines.push_back(l);
        s = e + 1;
    }
    return lines;
}

void widen(SoupMesh& m, const float* t)
{
    for (int c = 0; c < 3; ++c) {
        float lo = fminf(t[c], fminf(t[3 + c], t[6 + c]));
        float hi = fmaxf(t[c], fmaxf(t[3 + c], t[6 + c]));
        m.bbmin[c] = fminf(lo, m.bbmin[c]);
        m.bbmax[c] = fmaxf(hi, m.bbmax[c]);
    }
}

struct Material {
    std::string name;
    float kd[3] = {0.0f, 0.0f, 0.0f};
};

std::string load_mtl(const std::string& path, std::vector<Material>& mats, std::map<std::string, int>& index)
{
    std::string data;
    if (!read_file(path, data)) return "material library " + path + " could not be read";
    bool open = false;
    Material cur;
    auto flush = [&]() {
        if (open && !cur.name.empty()) {
            index[cur.name] = (int)mats.size();
            mats.push_back(cur);
        }
    };
    for (const std::string& line : split_lines(data)) {
        std::vector<std::string> w = words(line);
        if (w.empty()) continue;
        if (w[0] == "newmtl") {
            flush();
            cur = Material();
            cur.name = rest_of(line, "newmtl");
            open = true;
        } else if (w[0] == "Kd" && open) {
            if (w.size() < 4 || !parse_f32(w[1], cur.kd[0]) || !parse_f32(w[2], cur.kd[1]) ||
                !parse_f32(w[3], cur.kd[2]))
                return "material library " + path + ": bad Kd line";
        }
    }
    flush();
    return "";
}

std::string dir_of(const std::string& path)
{
    size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

std::string load_obj(const std::string& path, const std::string& data, std::vector<SoupMesh>& out)
{
    struct Model {
        std::string name;
        int mat;
        std::vector<uint32_t> corners;   // position index per triangle corner
    };
    std::vector<float> pos, vcol;
    std::vector<Model> models;
    std::vector<Material> mats;
    std::map<std::string, int> mat_index;
    std::string name = "unnamed";
    int mat = -1;
    std::vector<uint32_t> faces;
    // a model's faces are resolved when it ends, so they may name vertices
    // defined later inside the model
    bool bad_index = false;
    auto end_model = [&]() {
        if (!faces.empty()) {
            const uint32_t npos = (uint32_t)(pos.size() / 3);
            for (uint32_t v : faces) bad_index |= v >= npos;
            models.push_back(Model{name, mat, faces});
            faces.clear();
        }
    };
    size_t ln = 0;
    for (const std::string& line : split_lines(data)) {
        ++ln;
        std::vector<std::string> w = words(line);
        if (w.empty()) continue;
        const std::string& k = w[0];
        if (k == "v") {
            float p[6];
            const size_t nv = w.size() - 1;
            if (nv < 3 || !parse_f32(w[1], p[0]) || !parse_f32(w[2], p[1]) || !parse_f32(w[3], p[2]))
                return "line " + std::to_string(ln) + ": bad vertex position";
            pos.insert(pos.end(), p, p + 3);
            if (nv >= 6) {
                if (!parse_f32(w[4], p[3]) || !parse_f32(w[5], p[4]) || !parse_f32(w[6], p[5]))
                    return "line " + std::to_string(ln) + ": bad vertex colour";
                vcol.insert(vcol.end(), p + 3, p + 6);
            } else if (nv != 3) {
                return "line " + std::to_string(ln) + ": vertex with " + std::to_string(nv) +
                       " components (3 or x y z r g b)";
            }
        } else if (k == "f") {
            std::vector<uint32_t> ids;
            const long long npos = (long long)(pos.size() / 3);
            for (size_t i = 1; i < w.size(); ++i) {
                const std::string v = w[i].substr(0, w[i].find('/'));
                long long idx;
                if (!parse_int(v, idx) || idx == 0) return "line " + std::to_string(ln) + ": bad face index";
                idx = idx < 0 ? npos + idx : idx - 1;   // negative: relative to the vertices so far
                if (idx < 0 || idx >= (1ll << 32) - 1) return "line " + std::to_string(ln) + ": face index out of range";
                ids.push_back((uint32_t)idx);
            }
            // points and lines are ignored; polygons fan out from the first corner
            for (size_t i = 2; i < ids.size(); ++i) {
                faces.push_back(ids[0]);
                faces.push_back(ids[i - 1]);
                faces.push_back(ids[i]);
            }
        } else if (k == "o" || k == "g") {
            end_model();
            name = rest_of(line, k);
            if (name.empty()) name = "unnamed";
        } else if (k == "mtllib") {
            std::string err = load_mtl(dir_of(path) + rest_of(line, "mtllib"), mats, mat_index);
            if (!err.empty()) return err;
        } else if (k == "usemtl") {
            auto it = mat_index.find(rest_of(line, "usemtl"));
            if (it != mat_index.end()) {
                if (it->second != mat) end_model();
                mat = it->second;
            }
        }
    }
    end_model();
    if (bad_index) return "face index beyond the vertices of its model";
    const bool has_vcol = !vcol.empty();
    if (has_vcol && vcol.size() != pos.size()) return "vertex colours on some positions only";
    for (const Model& md : models) {
        SoupMesh m;
        m.name = md.name;
        const size_t n = md.corners.size() / 3;
        m.xyz.resize(n * 9);
        m.rgb.assign(n * 3, 1);
        for (int c = 0; c < 3; ++c) m.bbmin[c] = m.bbmax[c] = 0.0f;
        if (!mats.empty() && md.mat < 0) return "model " + md.name + " has no material";
        for (size_t t = 0; t < n; ++t) {
            for (int v = 0; v < 3; ++v)
                for (int c = 0; c < 3; ++c) m.xyz[t * 9 + v * 3 + c] = pos[(size_t)md.corners[t * 3 + v] * 3 + c];
            if (!mats.empty()) {
                const float* src = mats[md.mat].kd;
                if (has_vcol) src = &vcol[(size_t)md.corners[t * 3] * 3];
                for (int c = 0; c < 3; ++c) m.rgb[t * 3 + c] = as_u8(src[c] * 255.0f);
            }
            widen(m, &m.xyz[t * 9]);
        }
        out.push_back(std::move(m));
    }
    return "";
}

std::string load_stl(const std::string& data, std::vector<SoupMesh>& out)
{
    SoupMesh m;
    m.name = "stl";
    for (int c = 0; c < 3; ++c) {
        m.bbmin[c] = FLT_MAX;
        m.bbmax[c] = -FLT_MAX;
    }
    const size_t eol = data.find('\n');
    const std::string first = data.substr(0, eol);
    if (first.compare(0, 6, "solid ") == 0) {
        // ASCII: facet normal f f f / outer loop / vertex f f f x3 / endloop / endfacet ... endsolid
        std::istringstream is(eol == std::string::npos ? std::string() : data.substr(eol + 1));
        std::string t;
        auto expect = [&](const char* s) { return (is >> t) && t == s; };
        auto floats = [&](float* f, int k) {
            for (int i = 0; i < k; ++i)
                if (!(is >> t) || !parse_f32(t, f[i])) return false;
            return true;
        };
        while (is >> t) {
            if (t.compare(0, 8, "endsolid") == 0) break;
            float nrm[3], v[9];
            if (t != "facet" || !expect("normal") || !floats(nrm, 3) || !expect("outer") || !expect("loop"))
                return "ASCII STL: bad facet header";
            for (int k = 0; k < 3; ++k)
                if (!expect("vertex") || !floats(v + 3 * k, 3)) return "ASCII STL: bad vertex";
            if (!expect("endloop") || !expect("endfacet")) return "ASCII STL: bad facet end";
            m.xyz.insert(m.xyz.end(), v, v + 9);
        }
    } else {
        if (data.size() < 84) return "binary STL shorter than its header";
        uint32_t count;
        memcpy(&count, data.data() + 80, 4);
        if ((uint64_t)data.size() < 84 + 50ull * count) return "binary STL shorter than its facet count";
        m.xyz.resize((size_t)count * 9);
        for (uint32_t i = 0; i < count; ++i) memcpy(&m.xyz[(size_t)i * 9], data.data() + 84 + 50ull * i + 12, 36);
    }
    const size_t n = m.xyz.size() / 9;
    m.rgb.resize(n * 3);
    for (size_t i = 0; i < n; ++i) {
        m.rgb[i * 3 + 0] = 0xFF;
        m.rgb[i * 3 + 1] = 0xFF;
        m.rgb[i * 3 + 2] = 0x00;
        widen(m, &m.xyz[i * 9]);
    }
    out.push_back(std::move(m));
    return "";
}

}  // namespace

std::string load_model_file(const std::string& path, std::vector<SoupMesh>& out)
{
    const size_t dot = path.find_last_of('.');
    const size_t slash = path.find_last_of('/');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash))
        return "couldn't determine filename extension";
    std::string ext = path.substr(dot + 1);
    std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return (char)tolower(c); });
    if (ext != "obj" && ext != "stl") return "unknown filename extension";
    std::string data;
    if (!read_file(path, data)) return "could not read " + path;
    return ext == "obj" ? load_obj(path, data, out) : load_stl(data, out);
}

}  // namespace sloth
