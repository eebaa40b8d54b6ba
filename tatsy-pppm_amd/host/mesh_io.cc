// mesh_io.cc -- triangle mesh files for the GPU scene (pppm_mesh_load).
//
// Replaces Trimesh::load / loadPly / loadObj (sources/trimesh.cc:125-259).
// The reference reads binary little-endian PLY with exactly three float
// coordinates per vertex and triangles listed as uchar count + three ints; on
// those files this reader returns the same vertex and face arrays.  It also
// accepts what the reference rejects or misreads: ASCII and big-endian PLY,
// vertex elements with more properties (normals, colours) and double
// coordinates, other list types, polygons (fan-triangulated), and OBJ faces
// with 1-based, negative or slash-separated indices (the reference keeps OBJ
// indices 1-based, an off-by-one).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pppm.h"

namespace {

enum Ty { T_I8, T_U8, T_I16, T_U16, T_I32, T_U32, T_F32, T_F64, T_BAD };

Ty type_of(const std::string& s) {
    if (s == "char" || s == "int8") return T_I8;
    if (s == "uchar" || s == "uint8") return T_U8;
    if (s == "short" || s == "int16") return T_I16;
    if (s == "ushort" || s == "uint16") return T_U16;
    if (s == "int" || s == "int32") return T_I32;
    if (s == "uint" || s == "uint32") return T_U32;
    if (s == "float" || s == "float32") return T_F32;
    if (s == "double" || s == "float64") return T_F64;
    return T_BAD;
}

int size_of(Ty t) {
    switch (t) {
        case T_I8: case T_U8: return 1;
        case T_I16: case T_U16: return 2;
        case T_I32: case T_U32: case T_F32: return 4;
        case T_F64: return 8;
        default: return 0;
    }
}

struct Prop {
    std::string name;
    Ty type = T_BAD;
    bool list = false;
    Ty count_type = T_BAD;
};

struct Element {
    std::string name;
    int64_t count = 0;
    std::vector<Prop> props;
};

struct Reader {
    std::ifstream& in;
    int format;  // 0 ascii, 1 binary little endian, 2 binary big endian
    std::string err;

    bool value(Ty t, double* out) {
        if (format == 0) {
            std::string tok;
            if (!(in >> tok)) return fail("unexpected end of file");
            char* end = nullptr;
            errno = 0;
            *out = std::strtod(tok.c_str(), &end);
            if (end == tok.c_str() || *end != '\0') return fail("bad number '" + tok + "'");
            return true;
        }
        unsigned char b[8];
        const int n = size_of(t);
        if (!in.read(reinterpret_cast<char*>(b), n)) return fail("unexpected end of file");
        if (format == 2)
            for (int i = 0; i < n / 2; i++) std::swap(b[i], b[n - 1 - i]);
        switch (t) {
            case T_I8: { int8_t v; std::memcpy(&v, b, 1); *out = v; break; }
            case T_U8: { uint8_t v; std::memcpy(&v, b, 1); *out = v; break; }
            case T_I16: { int16_t v; std::memcpy(&v, b, 2); *out = v; break; }
            case T_U16: { uint16_t v; std::memcpy(&v, b, 2); *out = v; break; }
            case T_I32: { int32_t v; std::memcpy(&v, b, 4); *out = v; break; }
            case T_U32: { uint32_t v; std::memcpy(&v, b, 4); *out = v; break; }
            case T_F32: { float v; std::memcpy(&v, b, 4); *out = v; break; }
            case T_F64: { double v; std::memcpy(&v, b, 8); *out = v; break; }
            default: return fail("bad property type");
        }
        return true;
    }
    bool fail(const std::string& m) {
        if (err.empty()) err = m;
        return false;
    }
};

bool load_ply(const char* path, std::vector<float>* verts, std::vector<int32_t>* faces, std::string* err) {
    std::ifstream in(path, std::ios::in | std::ios::binary);
    if (!in.is_open()) return *err = std::string("cannot open ") + path, false;
    std::string line;
    std::getline(in, line);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line != "ply") return *err = "not a PLY file (no 'ply' magic)", false;
    int format = -1;
    std::vector<Element> elems;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "format") {
            std::string f;
            ss >> f;
            format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
            if (format < 0) return *err = "unknown PLY format " + f, false;
        } else if (key == "element") {
            Element e;
            ss >> e.name >> e.count;
            if (e.count < 0) return *err = "negative element count", false;
            elems.push_back(e);
        } else if (key == "property") {
            if (elems.empty()) return *err = "property before any element", false;
            Prop p;
            std::string t;
            ss >> t;
            if (t == "list") {
                std::string ct, it;
                ss >> ct >> it >> p.name;
                p.list = true;
                p.count_type = type_of(ct);
                p.type = type_of(it);
                if (p.count_type == T_BAD || p.type == T_BAD) return *err = "bad list property " + line, false;
            } else {
                p.type = type_of(t);
                ss >> p.name;
                if (p.type == T_BAD) return *err = "bad property " + line, false;
            }
            elems.back().props.push_back(p);
        } else if (key == "end_header") {
            break;
        }
    }
    if (format < 0) return *err = "PLY header has no format line", false;
    Reader rd{in, format, ""};
    for (const Element& e : elems) {
        const bool is_vertex = e.name == "vertex", is_face = e.name == "face";
        int ix = -1, iy = -1, iz = -1, iv = -1;
        for (size_t k = 0; k < e.props.size(); k++) {
            const std::string& n = e.props[k].name;
            if (n == "x") ix = (int)k;
            if (n == "y") iy = (int)k;
            if (n == "z") iz = (int)k;
            if ((n == "vertex_indices" || n == "vertex_index") && e.props[k].list) iv = (int)k;
        }
        if (is_vertex && (ix < 0 || iy < 0 || iz < 0)) return *err = "vertex element without x, y, z", false;
        if (is_face && iv < 0) return *err = "face element without a vertex_indices list", false;
        for (int64_t r = 0; r < e.count; r++) {
            double xyz[3] = {0, 0, 0};
            std::vector<double> poly;
            for (size_t k = 0; k < e.props.size(); k++) {
                const Prop& p = e.props[k];
                if (p.list) {
                    double cnt;
                    if (!rd.value(p.count_type, &cnt)) return *err = rd.err, false;
                    for (int64_t j = 0; j < (int64_t)cnt; j++) {
                        double v;
                        if (!rd.value(p.type, &v)) return *err = rd.err, false;
                        if ((int)k == iv) poly.push_back(v);
                    }
                } else {
                    double v;
                    if (!rd.value(p.type, &v)) return *err = rd.err, false;
                    if ((int)k == ix) xyz[0] = v;
                    if ((int)k == iy) xyz[1] = v;
                    if ((int)k == iz) xyz[2] = v;
                }
            }
            if (is_vertex)
                for (int d = 0; d < 3; d++) verts->push_back((float)xyz[d]);
            if (is_face) {
                if (poly.size() < 3) return *err = "face with fewer than three vertices", false;
                for (size_t j = 1; j + 1 < poly.size(); j++) {  // fan
                    faces->push_back((int32_t)poly[0]);
                    faces->push_back((int32_t)poly[j]);
                    faces->push_back((int32_t)poly[j + 1]);
                }
            }
        }
    }
    return true;
}

bool load_obj(const char* path, std::vector<float>* verts, std::vector<int32_t>* faces, std::string* err) {
    std::ifstream in(path);
    if (!in.is_open()) return *err = std::string("cannot open ") + path, false;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string typ;
        if (!(ss >> typ) || typ[0] == '#') continue;
        if (typ == "v") {
            double x, y, z;
            if (!(ss >> x >> y >> z)) return *err = "bad vertex line: " + line, false;
            verts->push_back((float)x);
            verts->push_back((float)y);
            verts->push_back((float)z);
        } else if (typ == "f") {
            std::vector<int32_t> poly;
            std::string tok;
            const int64_t nv = (int64_t)verts->size() / 3;
            while (ss >> tok) {
                const long i = std::strtol(tok.c_str(), nullptr, 10);  // index before any '/'
                if (i == 0) return *err = "bad face index in: " + line, false;
                poly.push_back((int32_t)(i > 0 ? i - 1 : nv + i));
            }
            if (poly.size() < 3) return *err = "face with fewer than three vertices: " + line, false;
            for (size_t j = 1; j + 1 < poly.size(); j++) {
                faces->push_back(poly[0]);
                faces->push_back(poly[j]);
                faces->push_back(poly[j + 1]);
            }
        }
        // vn, vt, o, g, s, usemtl, mtllib: not geometry
    }
    return true;
}

}  // namespace

extern "C" {

int pppm_mesh_load(const char* path, pppm_mesh* out, char* err, int32_t err_len) {
    if (!path || !out) return PPPM_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    std::vector<float> v;
    std::vector<int32_t> f;
    std::string e;
    const std::string p(path);
    const size_t dot = p.find_last_of('.');
    const std::string ext = dot == std::string::npos ? "" : p.substr(dot);
    bool ok;
    if (ext == ".ply" || ext == ".PLY") ok = load_ply(path, &v, &f, &e);
    else if (ext == ".obj" || ext == ".OBJ") ok = load_obj(path, &v, &f, &e);
    else ok = false, e = "mesh files are .ply or .obj";
    const int64_t nv = (int64_t)v.size() / 3;
    for (size_t k = 0; ok && k < f.size(); k++)
        if (f[k] < 0 || f[k] >= nv) ok = false, e = "face index out of range";
    if (!ok) {
        if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", e.c_str());
        return PPPM_ERR_INVALID;
    }
    out->n_vertices = nv;
    out->n_faces = (int64_t)f.size() / 3;
    out->vertices = static_cast<float*>(std::malloc(sizeof(float) * (v.size() ? v.size() : 1)));
    out->faces = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (f.size() ? f.size() : 1)));
    if (!out->vertices || !out->faces) {
        pppm_mesh_free(out);
        return PPPM_ERR_NOMEM;
    }
    std::memcpy(out->vertices, v.data(), sizeof(float) * v.size());
    std::memcpy(out->faces, f.data(), sizeof(int32_t) * f.size());
    return PPPM_OK;
}

void pppm_mesh_free(pppm_mesh* m) {
    if (!m) return;
    std::free(m->vertices);
    std::free(m->faces);
    std::memset(m, 0, sizeof(*m));
}

}  // extern "C"
