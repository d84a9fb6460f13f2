// pbrt-v4 subset loader (pbrt.h).
#include "pbrt.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "../csrc/mrt_device.h"

namespace mrt::pbrt {

namespace {

// ------------------------------------------------------------------------------------------
// tokens

struct Lexer {
    std::string text;
    size_t p = 0;
    bool next(std::string& tok, bool& quoted) {
        while (p < text.size()) {
            const char c = text[p];
            if (c == '#') {
                while (p < text.size() && text[p] != '\n') p++;
            } else if (std::isspace(static_cast<unsigned char>(c))) {
                p++;
            } else {
                break;
            }
        }
        if (p >= text.size()) return false;
        quoted = false;
        const char c = text[p];
        if (c == '"') {
            const size_t q = text.find('"', p + 1);
            if (q == std::string::npos) throw std::runtime_error("pbrt: unterminated string");
            tok = text.substr(p + 1, q - p - 1);
            p = q + 1;
            quoted = true;
        } else if (c == '[' || c == ']') {
            tok = std::string(1, c);
            p++;
        } else {
            const size_t s = p;
            while (p < text.size() && !std::isspace(static_cast<unsigned char>(text[p])) && text[p] != '[' &&
                   text[p] != ']' && text[p] != '"')
                p++;
            tok = text.substr(s, p - s);
        }
        return true;
    }
};

struct Param {
    std::string type, name;
    std::vector<double> nums;
    std::vector<std::string> strs;
    std::vector<bool> bools;
};
using Params = std::map<std::string, Param>;

float pf(const Params& ps, const std::string& n, float def) {
    auto it = ps.find(n);
    return it != ps.end() && !it->second.nums.empty() ? float(it->second.nums[0]) : def;
}
bool pb(const Params& ps, const std::string& n, bool def) {
    auto it = ps.find(n);
    return it != ps.end() && !it->second.bools.empty() ? bool(it->second.bools[0]) : def;
}
std::string ps_(const Params& ps, const std::string& n, const std::string& def) {
    auto it = ps.find(n);
    return it != ps.end() && !it->second.strs.empty() ? it->second.strs[0] : def;
}
bool prgb(const Params& ps, const std::string& n, float out[3]) {
    auto it = ps.find(n);
    if (it == ps.end()) return false;
    const Param& p = it->second;
    if (p.nums.size() >= 3) {
        for (int k = 0; k < 3; k++) out[k] = float(p.nums[k]);
        return true;
    }
    if (p.nums.size() == 1) { // a constant spectrum
        for (int k = 0; k < 3; k++) out[k] = float(p.nums[0]);
        return true;
    }
    return false;
}

// 4x4 row-major affine transforms
struct M4 {
    double m[16];
    static M4 id() {
        M4 r{};
        r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0;
        return r;
    }
    M4 operator*(const M4& b) const {
        M4 r{};
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                for (int k = 0; k < 4; k++) r.m[4 * i + j] += m[4 * i + k] * b.m[4 * k + j];
        return r;
    }
    void point(const float in[3], float out[3]) const {
        double o[3];
        for (int i = 0; i < 3; i++) o[i] = m[4 * i] * in[0] + m[4 * i + 1] * in[1] + m[4 * i + 2] * in[2] + m[4 * i + 3];
        for (int i = 0; i < 3; i++) out[i] = float(o[i]);
    }
    double det3() const {
        return m[0] * (m[5] * m[10] - m[6] * m[9]) - m[1] * (m[4] * m[10] - m[6] * m[8]) +
               m[2] * (m[4] * m[9] - m[5] * m[8]);
    }
    // inverse transpose of the upper 3x3 (normal transform)
    void normal(const float in[3], float out[3]) const {
        const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
        const double co[9] = {e * i - f * h, -(d * i - f * g), d * h - e * g, -(b * i - c * h), a * i - c * g,
                              -(a * h - b * g), b * f - c * e, -(a * f - c * d), a * e - b * d};
        // cofactor matrix = det * inverse transpose: the sign of det restores the orientation
        const double sgn = det3() < 0.0 ? -1.0 : 1.0;
        double o[3];
        for (int r = 0; r < 3; r++) o[r] = sgn * (co[3 * r] * in[0] + co[3 * r + 1] * in[1] + co[3 * r + 2] * in[2]);
        const double l = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
        for (int r = 0; r < 3; r++) out[r] = l > 1e-30 ? float(o[r] / l) : 0.f;
    }
    bool invert(M4& out) const {
        // Gauss-Jordan on a 4x4
        double a[4][8];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 8; j++) a[i][j] = j < 4 ? m[4 * i + j] : (j - 4 == i ? 1.0 : 0.0);
        for (int c = 0; c < 4; c++) {
            int piv = c;
            for (int r = c + 1; r < 4; r++)
                if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
            if (std::fabs(a[piv][c]) < 1e-300) return false;
            for (int j = 0; j < 8; j++) std::swap(a[c][j], a[piv][j]);
            const double inv = 1.0 / a[c][c];
            for (int j = 0; j < 8; j++) a[c][j] *= inv;
            for (int r = 0; r < 4; r++) {
                if (r == c) continue;
                const double f = a[r][c];
                for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
            }
        }
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) out.m[4 * i + j] = a[i][j + 4];
        return true;
    }
};

// ------------------------------------------------------------------------------------------
// PLY meshes (ascii, binary little / big endian; .gz through zlib)

struct PlyMesh {
    std::vector<float> pos, nrm, uv;
    std::vector<uint32_t> tri;
};

bool read_all(const std::string& path, std::string& out) {
    const bool gz = path.size() > 3 && path.compare(path.size() - 3, 3, ".gz") == 0;
    if (!gz) {
        std::ifstream in(path, std::ios::binary);
        if (!in) return false;
        out.assign((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        return true;
    }
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    char buf[1 << 16];
    int n;
    while ((n = gzread(f, buf, sizeof buf)) > 0) out.append(buf, size_t(n));
    const bool ok = n == 0;
    gzclose(f);
    return ok;
}

struct PlyProperty {
    std::string name;
    char type = 'f';      // c C s S i I f d: int8 uint8 int16 uint16 int32 uint32 float32 float64
    char count_type = 0;  // list properties: the type of the element count
};
struct PlyElement {
    std::string name;
    size_t count = 0;
    std::vector<PlyProperty> props;
};

char ply_type(const std::string& t) {
    static const std::map<std::string, char> types = {
        {"char", 'c'},  {"int8", 'c'},   {"uchar", 'C'}, {"uint8", 'C'},  {"short", 's'},   {"int16", 's'},
        {"ushort", 'S'}, {"uint16", 'S'}, {"int", 'i'},   {"int32", 'i'},  {"uint", 'I'},    {"uint32", 'I'},
        {"float", 'f'}, {"float32", 'f'}, {"double", 'd'}, {"float64", 'd'}};
    auto it = types.find(t);
    return it == types.end() ? 0 : it->second;
}
size_t ply_size(char t) { return t == 'c' || t == 'C' ? 1 : t == 's' || t == 'S' ? 2 : t == 'd' ? 8 : 4; }

// Values of the PLY body in file order.
struct PlyBody {
    const std::string& data;
    size_t p;
    int format; // 0 ascii, 1 binary little endian, 2 binary big endian
    bool failed = false;
    double next(char t) {
        if (format == 0) {
            while (p < data.size() && std::isspace(static_cast<unsigned char>(data[p]))) p++;
            const char* b = data.c_str() + p;
            char* e = nullptr;
            const double v = std::strtod(b, &e);
            if (e == b) failed = true;
            p += size_t(e - b);
            return v;
        }
        const size_t n = ply_size(t);
        if (p + n > data.size()) {
            failed = true;
            return 0.0;
        }
        unsigned char raw[8];
        std::memcpy(raw, data.data() + p, n);
        p += n;
        if (format == 2) std::reverse(raw, raw + n);
        switch (t) {
        case 'c': return double(int8_t(raw[0]));
        case 'C': return double(raw[0]);
        case 's': { int16_t v; std::memcpy(&v, raw, 2); return v; }
        case 'S': { uint16_t v; std::memcpy(&v, raw, 2); return v; }
        case 'i': { int32_t v; std::memcpy(&v, raw, 4); return v; }
        case 'I': { uint32_t v; std::memcpy(&v, raw, 4); return v; }
        case 'f': { float v; std::memcpy(&v, raw, 4); return v; }
        default: { double v; std::memcpy(&v, raw, 8); return v; }
        }
    }
};

bool read_ply(const std::string& path, PlyMesh& m, std::string& why) {
    std::string data;
    if (!read_all(path, data)) {
        why = "cannot read the file";
        return false;
    }
    std::istringstream header(data);
    std::string line;
    std::getline(header, line);
    if (line.compare(0, 3, "ply") != 0) {
        why = "not a PLY file";
        return false;
    }
    int format = -1;
    std::vector<PlyElement> elems;
    size_t body = std::string::npos;
    while (std::getline(header, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ls(line);
        std::string kw;
        ls >> kw;
        if (kw == "format") {
            std::string f;
            ls >> f;
            format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
        } else if (kw == "element") {
            PlyElement e;
            ls >> e.name >> e.count;
            elems.push_back(e);
        } else if (kw == "property" && !elems.empty()) {
            std::string t;
            ls >> t;
            PlyProperty pr;
            if (t == "list") {
                std::string ct, vt;
                ls >> ct >> vt >> pr.name;
                pr.count_type = ply_type(ct);
                pr.type = ply_type(vt);
                if (!pr.count_type) pr.type = 0;
            } else {
                pr.type = ply_type(t);
                ls >> pr.name;
            }
            if (!pr.type) {
                why = "unknown property type in '" + line + "'";
                return false;
            }
            elems.back().props.push_back(pr);
        } else if (kw == "end_header") {
            body = size_t(header.tellg());
            break;
        }
    }
    if (format < 0 || body == std::string::npos) {
        why = "no format or no end_header";
        return false;
    }
    PlyBody in{data, body, format};
    size_t nv = 0;
    bool got_vertices = false, got_faces = false;
    for (const PlyElement& e : elems) {
        const bool is_vertex = e.name == "vertex", is_face = e.name == "face";
        int ip[3] = {-1, -1, -1}, in_[3] = {-1, -1, -1}, it[2] = {-1, -1}, iv = -1;
        for (size_t k = 0; k < e.props.size(); k++) {
            const std::string& n = e.props[k].name;
            const int kk = int(k);
            if (n == "x") ip[0] = kk; else if (n == "y") ip[1] = kk; else if (n == "z") ip[2] = kk;
            else if (n == "nx") in_[0] = kk; else if (n == "ny") in_[1] = kk; else if (n == "nz") in_[2] = kk;
            else if (n == "u" || n == "s" || n == "texture_u" || n == "texture_s") it[0] = kk;
            else if (n == "v" || n == "t" || n == "texture_v" || n == "texture_t") it[1] = kk;
            else if (n == "vertex_indices" || n == "vertex_index") iv = kk;
        }
        if (is_vertex && (ip[0] < 0 || ip[1] < 0 || ip[2] < 0)) {
            why = "vertices without x, y, z";
            return false;
        }
        const bool with_n = in_[0] >= 0 && in_[1] >= 0 && in_[2] >= 0, with_uv = it[0] >= 0 && it[1] >= 0;
        std::vector<double> row(e.props.size());
        std::vector<uint32_t> poly;
        for (size_t r = 0; r < e.count && !in.failed; r++) {
            poly.clear();
            for (size_t k = 0; k < e.props.size(); k++) {
                const PlyProperty& pr = e.props[k];
                if (pr.count_type) {
                    const size_t cnt = size_t(in.next(pr.count_type));
                    for (size_t q = 0; q < cnt && !in.failed; q++) {
                        const double v = in.next(pr.type);
                        if (is_face && int(k) == iv) poly.push_back(uint32_t(v));
                    }
                } else {
                    row[k] = in.next(pr.type);
                }
            }
            if (is_vertex) {
                for (int a = 0; a < 3; a++) m.pos.push_back(float(row[size_t(ip[a])]));
                if (with_n)
                    for (int a = 0; a < 3; a++) m.nrm.push_back(float(row[size_t(in_[a])]));
                if (with_uv)
                    for (int a = 0; a < 2; a++) m.uv.push_back(float(row[size_t(it[a])]));
            } else if (is_face) {
                for (size_t q = 1; q + 1 < poly.size(); q++) // polygons as triangle fans
                    if (poly[0] < nv && poly[q] < nv && poly[q + 1] < nv) m.tri.insert(m.tri.end(), {poly[0], poly[q], poly[q + 1]});
            }
        }
        if (in.failed) {
            why = "the body ends early or holds a malformed value (element '" + e.name + "')";
            return false;
        }
        if (is_vertex) {
            nv = e.count;
            got_vertices = true;
        }
        if (is_face) {
            if (!got_vertices || iv < 0) {
                why = "faces without vertex_indices, or before the vertices";
                return false;
            }
            got_faces = true;
        }
    }
    if (!got_vertices || !got_faces) {
        why = "no vertex or no face element";
        return false;
    }
    return true;
}

void normalize3(double v[3]) {
    const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0)
        for (int k = 0; k < 3; k++) v[k] /= l;
}

M4 look_at(const double e[3], const double t[3], const double u_in[3]) {
    // pbrt's LookAt: camera space looks down +z, x = normalize(cross(up, dir)), y = cross(dir, x)
    double dir[3] = {t[0] - e[0], t[1] - e[1], t[2] - e[2]};
    normalize3(dir);
    double up[3] = {u_in[0], u_in[1], u_in[2]};
    normalize3(up);
    double x[3] = {up[1] * dir[2] - up[2] * dir[1], up[2] * dir[0] - up[0] * dir[2], up[0] * dir[1] - up[1] * dir[0]};
    normalize3(x);
    const double y[3] = {dir[1] * x[2] - dir[2] * x[1], dir[2] * x[0] - dir[0] * x[2], dir[0] * x[1] - dir[1] * x[0]};
    M4 cam_to_world = M4::id();
    for (int r = 0; r < 3; r++) {
        cam_to_world.m[4 * r + 0] = x[r];
        cam_to_world.m[4 * r + 1] = y[r];
        cam_to_world.m[4 * r + 2] = dir[r];
        cam_to_world.m[4 * r + 3] = e[r];
    }
    M4 w2c;
    cam_to_world.invert(w2c);
    return w2c;
}

mrt_material openpbr_defaults() {
    mrt_material m{};
    m.model = MRT_MAT_OPENPBR;
    for (int k = 0; k < 3; k++) {
        m.base_color[k] = 1.f;
        m.sheen_color[k] = 1.f;
        m.transmission_color[k] = 1.f;
    }
    m.opacity = 1.f;
    m.specular_alpha[0] = m.specular_alpha[1] = 1.f;
    m.specular_weight = 1.f;
    m.specular_ior = 1.5f;
    m.coat_ior = 1.6f;
    m.sheen_alpha = 0.09f;
    return m;
}

struct Material {
    std::string type;
    Params params;
};

struct AreaLight {
    bool on = false;
    float L[3] = {1, 1, 1};
    float scale = 1.f;
    bool twosided = false;
};

struct State {
    M4 ctm = M4::id();
    int material = -1; // index into the material descriptions, -1 = pbrt's default
    AreaLight area;
    bool reverse = false;
};

struct Loader {
    Scene& sc;
    std::vector<Material> mats;
    std::map<std::string, int> named;
    std::map<std::string, uint32_t> converted; // material description -> scene material index
    std::vector<State> stack;
    State st;
    M4 world_to_camera = M4::id();
    double fov_deg = 90.0;
    bool have_camera = false;
    uint32_t film_w = 1280, film_h = 720;
    std::map<std::string, bool> warned;

    explicit Loader(Scene& s) : sc(s) {}

    void warn(const std::string& key, const std::string& msg) {
        if (warned[key]) return;
        warned[key] = true;
        sc.warnings.push_back(msg);
    }

    // roughness -> GGX alpha (pbrt_scene.cpp:332-364): sqrt(r) unless remaproughness is false
    void roughness(const Params& p, const std::string& prefix, float alpha[2]) {
        const bool remap = pb(p, "remaproughness", true);
        auto to_alpha = [&](float r) { return remap ? std::sqrt(r) : r; };
        const float v = pf(p, prefix + "roughness", 0.f);
        alpha[0] = alpha[1] = to_alpha(v);
        if (p.count(prefix + "uroughness") || p.count(prefix + "vroughness")) {
            alpha[0] = to_alpha(pf(p, prefix + "uroughness", v));
            alpha[1] = to_alpha(pf(p, prefix + "vroughness", v));
        }
    }

    // material description -> (OpenPBR entry, forced two-sided)
    mrt_material convert(int idx, bool& two_sided, int depth = 0) {
        mrt_material m = openpbr_defaults();
        two_sided = false;
        if (idx < 0 || depth > 4) { // pbrt's default: coateddiffuse, reflectance 0.5
            for (int k = 0; k < 3; k++) m.base_color[k] = 0.5f;
            m.specular_alpha[0] = m.specular_alpha[1] = 0.f;
            return m;
        }
        const Material& d = mats[size_t(idx)];
        const Params& p = d.params;
        const std::string& t = d.type;
        if (t == "diffuse" || t == "diffusetransmission" || t == "subsurface" || t == "hair" || t == "measured") {
            float c[3] = {0.5f, 0.5f, 0.5f};
            prgb(p, "reflectance", c);
            std::memcpy(m.base_color, c, 12);
            m.specular_weight = 0.f;
            m.specular_alpha[0] = m.specular_alpha[1] = 1.f;
            if (t != "diffuse") warn("mat_" + t, "pbrt material \"" + t + "\": loaded as a Lambert surface with its reflectance");
        } else if (t == "coateddiffuse") {
            float c[3] = {0.5f, 0.5f, 0.5f};
            prgb(p, "reflectance", c);
            std::memcpy(m.base_color, c, 12);
            m.specular_ior = pf(p, "eta", 1.5f);
            roughness(p, "", m.specular_alpha);
        } else if (t == "conductor" || t == "coatedconductor") {
            const std::string pre = t == "coatedconductor" ? "conductor." : "";
            m.metalness = 1.f;
            float c[3] = {0.955f, 0.637f, 0.538f}; // pbrt's default conductor (copper)
            if (p.count("reflectance")) {
                c[0] = c[1] = c[2] = 0.9f;
                prgb(p, "reflectance", c);
            } else if (p.count(pre + "eta") || p.count(pre + "k")) {
                float eta[3] = {0.2f, 0.2f, 0.2f}, k[3] = {3.9f, 3.9f, 3.9f};
                prgb(p, pre + "eta", eta);
                prgb(p, pre + "k", k);
                for (int a = 0; a < 3; a++)
                    c[a] = ((eta[a] - 1.f) * (eta[a] - 1.f) + k[a] * k[a]) / ((eta[a] + 1.f) * (eta[a] + 1.f) + k[a] * k[a]);
            }
            std::memcpy(m.base_color, c, 12);
            roughness(p, pre, m.specular_alpha);
            if (t == "coatedconductor") warn("mat_cc", "coatedconductor: only the conductor base is loaded, its dielectric coat is dropped");
        } else if (t == "dielectric" || t == "thindielectric") {
            m.specular_ior = pf(p, "eta", 1.5f);
            m.features = MRT_PBR_TRANSMISSION;
            m.transmission_weight = 1.f;
            for (int a = 0; a < 3; a++) m.transmission_color[a] = 1.f;
            roughness(p, "", m.specular_alpha);
            two_sided = true;
            if (t == "thindielectric") warn("mat_thin", "thindielectric: loaded as a solid (refracting) dielectric");
        } else if (t == "mix") {
            auto it = p.find("materials");
            if (it != p.end() && it->second.strs.size() == 2) {
                const std::string& chosen = it->second.strs[pf(p, "amount", 0.5f) >= 0.5f ? 1 : 0];
                auto n = named.find(chosen);
                if (n != named.end()) return convert(n->second, two_sided, depth + 1);
            }
            for (int k = 0; k < 3; k++) m.base_color[k] = 0.5f;
            m.specular_weight = 0.f;
            m.specular_alpha[0] = m.specular_alpha[1] = 1.f;
        } else {
            warn("mat_" + t, "material '" + t + "' is unsupported, using diffuse");
            for (int k = 0; k < 3; k++) m.base_color[k] = 0.5f;
            m.specular_weight = 0.f;
            m.specular_alpha[0] = m.specular_alpha[1] = 1.f;
        }
        return m;
    }

    // Shape "trianglemesh" (and "loopsubdiv", loaded unsubdivided): P, indices (optional for a
    // single triangle), N, uv or st (pbrt_scene.cpp:533-575).
    void trianglemesh(const Params& p) {
        auto P = p.find("P");
        if (P == p.end() || P->second.nums.size() < 9) {
            warn("tm_invalid", "a trianglemesh without at least one triangle of \"P\" was skipped");
            return;
        }
        const size_t nv = P->second.nums.size() / 3;
        std::vector<uint32_t> tri;
        auto I = p.find("indices");
        if (I != p.end()) {
            for (size_t k = 0; k + 2 < I->second.nums.size(); k += 3) {
                const uint32_t a = uint32_t(I->second.nums[k]), b = uint32_t(I->second.nums[k + 1]),
                               c = uint32_t(I->second.nums[k + 2]);
                if (a < nv && b < nv && c < nv) tri.insert(tri.end(), {a, b, c});
            }
        } else if (nv == 3) {
            tri = {0, 1, 2};
        } else {
            warn("tm_invalid", "a trianglemesh without \"indices\" and more than three points was skipped");
            return;
        }
        std::vector<float> pos(3 * nv), nrm, uv;
        for (size_t k = 0; k < 3 * nv; k++) pos[k] = float(P->second.nums[k]);
        auto N = p.find("N");
        if (N != p.end() && N->second.nums.size() == 3 * nv) {
            nrm.resize(3 * nv);
            for (size_t k = 0; k < 3 * nv; k++) nrm[k] = float(N->second.nums[k]);
        }
        auto UV = p.find("uv");
        if (UV == p.end()) UV = p.find("st");
        if (UV != p.end() && UV->second.nums.size() >= 2 * nv) {
            uv.resize(2 * nv);
            for (size_t k = 0; k < 2 * nv; k++) uv[k] = float(UV->second.nums[k]);
        }
        emit_mesh(pos, nrm, uv, tri);
    }

    // Shape "sphere": radius, zmin, zmax, phimax tessellated into 64 slices x 32 stacks with
    // analytic normals (pbrt_scene.cpp:692-733).
    void sphere(const Params& p) {
        const float r = pf(p, "radius", 1.f);
        const float zmin = std::clamp(pf(p, "zmin", -r), -r, r), zmax = std::clamp(pf(p, "zmax", r), -r, r);
        const float phi_max = std::clamp(pf(p, "phimax", 360.f), 0.f, 360.f) * (float(M_PI) / 180.f);
        const float th0 = std::acos(std::clamp(zmax / r, -1.f, 1.f)), th1 = std::acos(std::clamp(zmin / r, -1.f, 1.f));
        constexpr int S = 64, T = 32;
        std::vector<float> pos, nrm, uv;
        for (int j = 0; j <= T; j++) {
            const float theta = th0 + (th1 - th0) * float(j) / T;
            for (int i = 0; i <= S; i++) {
                const float phi = phi_max * float(i) / S;
                const float n[3] = {std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)};
                for (int k = 0; k < 3; k++) {
                    pos.push_back(r * n[k]);
                    nrm.push_back(n[k]);
                }
                uv.push_back(float(i) / S);
                uv.push_back(1.f - float(j) / T);
            }
        }
        std::vector<uint32_t> tri;
        for (int j = 0; j < T; j++)
            for (int i = 0; i < S; i++) {
                const uint32_t a = uint32_t(j * (S + 1) + i), b = a + S + 1;
                tri.insert(tri.end(), {a, b, a + 1, a + 1, b, b + 1});
            }
        emit_mesh(pos, nrm, uv, tri);
    }

    // Shape "disk": height, radius, innerradius, phimax as a ring of 64 slices (pbrt_scene.cpp:735-765).
    void disk(const Params& p) {
        const float h = pf(p, "height", 0.f), r = pf(p, "radius", 1.f);
        const float ri = std::clamp(pf(p, "innerradius", 0.f), 0.f, r);
        const float phi_max = std::clamp(pf(p, "phimax", 360.f), 0.f, 360.f) * (float(M_PI) / 180.f);
        constexpr int S = 64;
        std::vector<float> pos, nrm, uv;
        for (int i = 0; i <= S; i++) {
            const float phi = phi_max * float(i) / S;
            const float dx = std::cos(phi), dy = std::sin(phi);
            pos.insert(pos.end(), {r * dx, r * dy, h, ri * dx, ri * dy, h});
            nrm.insert(nrm.end(), {0.f, 0.f, 1.f, 0.f, 0.f, 1.f});
            uv.insert(uv.end(), {float(i) / S, 0.f, float(i) / S, 1.f});
        }
        std::vector<uint32_t> tri;
        for (uint32_t i = 0; i < uint32_t(S); i++) {
            const uint32_t a = 2 * i;
            tri.insert(tri.end(), {a, a + 2, a + 1, a + 1, a + 2, a + 3});
        }
        emit_mesh(pos, nrm, uv, tri);
    }

    // Shape "plymesh": a PLY file (ascii or binary, optionally gzip-compressed) relative to the
    // scene file's directory (pbrt_scene.cpp:577-690,790-798).
    void plymesh(const Params& p, const std::string& base) {
        const std::string file = ps_(p, "filename", "");
        if (file.empty()) {
            warn("ply_nofile", "a plymesh without \"filename\" was skipped");
            return;
        }
        const std::string path = file[0] == '/' ? file : base + "/" + file;
        PlyMesh m;
        std::string why;
        if (!read_ply(path, m, why)) {
            warn("ply_" + path, "plymesh '" + path + "' was skipped: " + why);
            return;
        }
        emit_mesh(m.pos, m.nrm, m.uv, m.tri);
    }

    // Adds one shape: object-space positions, normals (empty: area-weighted from the positions,
    // pbrt_scene.cpp:55-70) and uvs (pbrt's convention, origin bottom-left; empty: zero), placed
    // by the current transform with the current material and area light (pbrt_scene.cpp:72-86,
    // 776-830). Positions are baked into world space.
    void emit_mesh(const std::vector<float>& pos, std::vector<float> nrm, const std::vector<float>& uv,
                   const std::vector<uint32_t>& tri) {
        if (tri.empty()) return;
        const size_t nv = pos.size() / 3;
        if (nrm.size() != 3 * nv) {
            nrm.assign(3 * nv, 0.f);
            for (size_t t = 0; t < tri.size(); t += 3) { // unnormalised face cross products
                const float* a = &pos[3 * tri[t]];
                const float* b = &pos[3 * tri[t + 1]];
                const float* c = &pos[3 * tri[t + 2]];
                const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                const float f[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                    e1[0] * e2[1] - e1[1] * e2[0]};
                for (int k = 0; k < 3; k++)
                    for (int a2 = 0; a2 < 3; a2++) nrm[3 * tri[t + k] + a2] += f[a2];
            }
        }
        const uint32_t vbase = uint32_t(sc.vertices.size());
        for (size_t v = 0; v < nv; v++) {
            mrt_vertex vx{};
            st.ctm.point(&pos[3 * v], vx.position);
            float on[3] = {nrm[3 * v], nrm[3 * v + 1], nrm[3 * v + 2]};
            const float ol = std::sqrt(on[0] * on[0] + on[1] * on[1] + on[2] * on[2]);
            float wn[3] = {0.f, 1.f, 0.f};
            if (ol > 1e-20f) {
                for (int k = 0; k < 3; k++) on[k] /= ol;
                st.ctm.normal(on, wn);
            }
            f3 n{wn[0], wn[1], wn[2]};
            const float l = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
            n = l > 1e-20f ? f3{n.x / l, n.y / l, n.z / l} : f3{0.f, 1.f, 0.f};
            vx.encoded_normal = encode_normal(n);
            if (uv.size() >= 2 * nv) { // pbrt's uv origin is bottom-left, the texture's top-left
                vx.uv[0] = float_to_half(uv[2 * v]);
                vx.uv[1] = float_to_half(1.f - uv[2 * v + 1]);
            }
            sc.vertices.push_back(vx);
        }
        // material (+ emission of an area light)
        bool two_sided_mat = false;
        mrt_material m = convert(st.material, two_sided_mat);
        if (st.area.on)
            for (int k = 0; k < 3; k++) m.emission[k] = st.area.L[k] * st.area.scale;
        const uint32_t mid = uint32_t(sc.materials.size());
        sc.materials.push_back(m);
        uint16_t flags = MRT_GEOM_PRETRANSFORMED | MRT_GEOM_INDEX_UINT32;
        const bool two_sided_emitter = st.area.on && st.area.twosided;
        auto add = [&](const std::vector<uint32_t>& idx, bool flip) {
            mrt_geometry g{};
            g.vertex_offset = vbase;
            g.index_offset = uint32_t(sc.indices.size());
            g.primitive_count = uint32_t(idx.size() / 3);
            g.material_id = uint16_t(mid);
            g.instance_mask = 0xFFu;
            uint16_t f = flags;
            if (flip) f |= MRT_GEOM_FLIP_FACING;
            g.flags = f;
            sc.indices.insert(sc.indices.end(), idx.begin(), idx.end());
            sc.geometries.push_back(g);
        };
        const bool flip = st.reverse != (st.ctm.det3() < 0.0);
        if (two_sided_emitter && !two_sided_mat) {
            // two single-sided sheets with opposite winding (both culled from behind)
            add(tri, flip);
            std::vector<uint32_t> rev(tri.size());
            for (size_t t = 0; t < tri.size(); t += 3) {
                rev[t] = tri[t];
                rev[t + 1] = tri[t + 2];
                rev[t + 2] = tri[t + 1];
            }
            add(rev, flip);
        } else {
            flags |= MRT_GEOM_TWO_SIDED;
            add(tri, flip);
        }
    }

    void infinite(const Params& p, const std::string& base) {
        if (!sc.env_rgba.empty()) {
            warn("env_multi", "multiple infinite lights, using the first");
            return;
        }
        sc.env_scale = pf(p, "scale", 1.f);
        const std::string file = ps_(p, "filename", "");
        if (!file.empty()) {
            const std::string path = file[0] == '/' ? file : base + "/" + file;
            if (!read_pfm(path, sc.env_rgba, sc.env_width, sc.env_height)) {
                warn("env_file", "environment image '" + path + "' not loaded (PFM only)");
                sc.env_rgba.clear();
                return;
            }
        } else {
            float L[3] = {1, 1, 1};
            prgb(p, "L", L);
            sc.env_rgba = {L[0], L[1], L[2], 1.f};
            sc.env_width = sc.env_height = 1;
        }
        // the image is evaluated at inverse(light CTM) * world direction
        M4 inv;
        if (st.ctm.invert(inv))
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) sc.env_to_local[3 * r + c] = float(inv.m[4 * r + c]);
    }

    void parse_file(const std::string& path) {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("pbrt: cannot open '" + path + "'");
        std::stringstream ss;
        ss << in.rdbuf();
        const size_t slash = path.find_last_of('/');
        const std::string base = slash == std::string::npos ? "." : path.substr(0, slash);
        Lexer lx{ss.str()};
        std::string tok;
        bool quoted;
        bool have = lx.next(tok, quoted);
        // n numbers starting at the current token (optionally bracketed)
        auto nums = [&](size_t n) {
            std::vector<double> v;
            while (v.size() < n) {
                if (!have) throw std::runtime_error("pbrt: unexpected end of file");
                if (tok != "[" && tok != "]") v.push_back(std::stod(tok));
                have = lx.next(tok, quoted);
            }
            if (have && tok == "]") have = lx.next(tok, quoted);
            return v;
        };
        // parameter list following a statement: "type name" value | [values]
        auto params = [&]() {
            Params ps;
            while (have && quoted) {
                std::istringstream decl(tok);
                Param p;
                decl >> p.type >> p.name;
                if (p.name.empty()) break; // a plain string argument, not a parameter
                have = lx.next(tok, quoted);
                auto value = [&](const std::string& v, bool q) {
                    if (p.type == "string" || p.type == "texture" || p.type == "spectrum" && q) {
                        p.strs.push_back(v);
                    } else if (p.type == "bool") {
                        p.bools.push_back(v == "true");
                    } else {
                        p.nums.push_back(std::stod(v));
                    }
                };
                if (have && tok == "[") {
                    while ((have = lx.next(tok, quoted)) && tok != "]") value(tok, quoted);
                    have = lx.next(tok, quoted);
                } else if (have) {
                    value(tok, quoted);
                    have = lx.next(tok, quoted);
                }
                ps[p.name] = std::move(p);
            }
            return ps;
        };
        auto string_arg = [&]() {
            if (!have || !quoted) throw std::runtime_error("pbrt: expected a string after a statement");
            const std::string s = tok;
            have = lx.next(tok, quoted);
            return s;
        };
        while (have) {
            const std::string stmt = tok;
            have = lx.next(tok, quoted);
            if (stmt == "LookAt") {
                const std::vector<double> v = nums(9);
                st.ctm = st.ctm * look_at(&v[0], &v[3], &v[6]);
            } else if (stmt == "Translate") {
                const std::vector<double> v = nums(3);
                M4 t = M4::id();
                t.m[3] = v[0];
                t.m[7] = v[1];
                t.m[11] = v[2];
                st.ctm = st.ctm * t;
            } else if (stmt == "Scale") {
                const std::vector<double> v = nums(3);
                M4 t = M4::id();
                t.m[0] = v[0];
                t.m[5] = v[1];
                t.m[10] = v[2];
                st.ctm = st.ctm * t;
            } else if (stmt == "Rotate") {
                std::vector<double> v = nums(4);
                double ax[3] = {v[1], v[2], v[3]};
                normalize3(ax);
                const double a = v[0] * M_PI / 180.0, s = std::sin(a), c = std::cos(a);
                M4 r = M4::id();
                const double x = ax[0], y = ax[1], z = ax[2];
                r.m[0] = x * x + (1 - x * x) * c;
                r.m[1] = x * y * (1 - c) - z * s;
                r.m[2] = x * z * (1 - c) + y * s;
                r.m[4] = x * y * (1 - c) + z * s;
                r.m[5] = y * y + (1 - y * y) * c;
                r.m[6] = y * z * (1 - c) - x * s;
                r.m[8] = x * z * (1 - c) - y * s;
                r.m[9] = y * z * (1 - c) + x * s;
                r.m[10] = z * z + (1 - z * z) * c;
                st.ctm = st.ctm * r;
            } else if (stmt == "Transform" || stmt == "ConcatTransform") {
                const std::vector<double> v = nums(16); // column-major in the file
                M4 t{};
                for (int r = 0; r < 4; r++)
                    for (int c = 0; c < 4; c++) t.m[4 * r + c] = v[size_t(4 * c + r)];
                st.ctm = stmt == "Transform" ? t : st.ctm * t;
            } else if (stmt == "Identity") {
                st.ctm = M4::id();
            } else if (stmt == "Camera") {
                const std::string type = string_arg();
                const Params p = params();
                if (type != "perspective") warn("cam_" + type, "camera '" + type + "' treated as perspective");
                fov_deg = pf(p, "fov", 90.f);
                world_to_camera = st.ctm;
                have_camera = true;
                st.ctm = M4::id();
            } else if (stmt == "Film") {
                string_arg();
                const Params p = params();
                film_w = uint32_t(pf(p, "xresolution", 1280.f));
                film_h = uint32_t(pf(p, "yresolution", 720.f));
            } else if (stmt == "WorldBegin") {
                st.ctm = M4::id();
            } else if (stmt == "AttributeBegin" || stmt == "TransformBegin") {
                stack.push_back(st);
            } else if (stmt == "AttributeEnd" || stmt == "TransformEnd") {
                if (stack.empty()) throw std::runtime_error("pbrt: unbalanced " + stmt);
                const State top = stack.back();
                stack.pop_back();
                if (stmt == "TransformEnd") {
                    st.ctm = top.ctm;
                } else {
                    st = top;
                }
            } else if (stmt == "ReverseOrientation") {
                st.reverse = !st.reverse;
            } else if (stmt == "Material") {
                Material m;
                m.type = string_arg();
                m.params = params();
                mats.push_back(std::move(m));
                st.material = int(mats.size()) - 1;
            } else if (stmt == "MakeNamedMaterial") {
                const std::string name = string_arg();
                Material m;
                m.params = params();
                m.type = ps_(m.params, "type", "diffuse");
                mats.push_back(std::move(m));
                named[name] = int(mats.size()) - 1;
            } else if (stmt == "NamedMaterial") {
                const std::string name = string_arg();
                auto it = named.find(name);
                if (it == named.end()) {
                    warn("nm_" + name, "unknown named material '" + name + "'");
                } else {
                    st.material = it->second;
                }
            } else if (stmt == "AreaLightSource") {
                string_arg();
                const Params p = params();
                st.area.on = true;
                prgb(p, "L", st.area.L);
                st.area.scale = pf(p, "scale", 1.f);
                st.area.twosided = pb(p, "twosided", false);
            } else if (stmt == "LightSource") {
                const std::string type = string_arg();
                const Params p = params();
                if (type == "infinite") {
                    infinite(p, base);
                } else {
                    warn("light_" + type, "light '" + type + "' is unsupported, skipping");
                }
            } else if (stmt == "Shape") {
                const std::string type = string_arg();
                const Params p = params();
                if (type == "trianglemesh" || type == "loopsubdiv") {
                    if (type == "loopsubdiv") warn("loopsubdiv", "loopsubdiv control meshes are loaded unsubdivided");
                    trianglemesh(p);
                } else if (type == "sphere") {
                    sphere(p);
                } else if (type == "disk") {
                    disk(p);
                } else if (type == "plymesh") {
                    plymesh(p, base);
                } else {
                    warn("shape_" + type, "shape '" + type + "' is unsupported, skipping");
                }
            } else if (stmt == "Include" || stmt == "Import") {
                const std::string f = string_arg();
                parse_file(f[0] == '/' ? f : base + "/" + f);
            } else if (stmt == "Texture") {
                string_arg();
                string_arg();
                string_arg();
                params();
                warn("texture", "textures are unsupported, skipping");
            } else if (stmt == "ObjectBegin" || stmt == "ObjectInstance") {
                throw std::runtime_error("pbrt: object instancing is not supported");
            } else if (stmt == "ObjectEnd" || stmt == "WorldEnd") {
            } else if (stmt == "Sampler" || stmt == "Integrator" || stmt == "PixelFilter" || stmt == "Accelerator" ||
                       stmt == "ColorSpace" || stmt == "Option" || stmt == "MakeNamedMedium" ||
                       stmt == "MediumInterface" || stmt == "CoordinateSystem" || stmt == "CoordSysTransform" ||
                       stmt == "Attribute") {
                while (have && (quoted || tok == "[" || tok == "]" ||
                                (!tok.empty() && (std::isdigit(static_cast<unsigned char>(tok[0])) || tok[0] == '-' ||
                                                  tok[0] == '.' || tok == "true" || tok == "false"))))
                    have = lx.next(tok, quoted);
            } else {
                throw std::runtime_error("pbrt: unknown statement '" + stmt + "'");
            }
        }
    }

    void finish() {
        Camera& c = sc.camera;
        c.width = film_w;
        c.height = film_h;
        if (have_camera) {
            M4 c2w;
            world_to_camera.invert(c2w);
            double fwd[3] = {c2w.m[2], c2w.m[6], c2w.m[10]}, up[3] = {c2w.m[1], c2w.m[5], c2w.m[9]};
            normalize3(fwd);
            normalize3(up);
            for (int k = 0; k < 3; k++) {
                c.eye[k] = float(c2w.m[4 * k + 3]);
                c.forward[k] = float(fwd[k]);
                c.up[k] = float(up[k]);
            }
            const double aspect = double(film_w) / double(film_h);
            const double fov = fov_deg * M_PI / 180.0;
            c.vfov = float(film_w >= film_h ? fov : 2.0 * std::atan(std::tan(0.5 * fov) / aspect));
            c.present = true;
        } else if (!sc.vertices.empty()) {
            default_camera(c);
        }
        if (sc.materials.empty()) sc.materials.push_back(openpbr_defaults());
    }

    // No camera in the file: a camera at (3, 3, 3) looking at the origin, 60 degrees, 16:9, then
    // fitted to the scene's bounds (pbrt_scene.cpp:994-1003) as Camera::look_at_bounding_box does
    // (camera.cpp:373-423): start above the box looking at its centre, then three times pan so the
    // projected box is centred and back off until all eight corners are inside 85 % of the field
    // of view. Float arithmetic, as the reference camera's.
    void default_camera(Camera& c) {
        float lo[3], hi[3];
        for (int k = 0; k < 3; k++) lo[k] = hi[k] = sc.vertices[0].position[k];
        for (const mrt_vertex& v : sc.vertices)
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], v.position[k]);
                hi[k] = std::max(hi[k], v.position[k]);
            }
        const f3 up{0.f, 1.f, 0.f};
        const float vfov = float(60.0 * M_PI / 180.0), aspect = 1920.f / 1080.f;
        const float tan_hy = std::tan(vfov * 0.85f * 0.5f), tan_hx = tan_hy * aspect;
        const f3 centre{(lo[0] + hi[0]) * 0.5f, (lo[1] + hi[1]) * 0.5f, (lo[2] + hi[2]) * 0.5f};
        f3 pos = f3{1.3f, 1.3f, 1.3f} * (hi[1] + 1e-3f), target = centre;
        const auto corner = [&](int i) { return f3{(i & 1) ? hi[0] : lo[0], (i & 2) ? hi[1] : lo[1], (i & 4) ? hi[2] : lo[2]}; };
        for (int iter = 0; iter < 3; iter++) {
            const f3 fwd = normalize(target - pos);
            const f3 r = normalize(cross(fwd, up));
            const f3 u = cross(r, fwd);
            float px0 = 3.4e38f, px1 = -3.4e38f, py0 = 3.4e38f, py1 = -3.4e38f;
            for (int i = 0; i < 8; i++) {
                const f3 v = corner(i) - pos;
                const float d = dot(v, fwd), px = dot(v, r) / d, py = dot(v, u) / d;
                px0 = std::min(px0, px);
                px1 = std::max(px1, px);
                py0 = std::min(py0, py);
                py1 = std::max(py1, py);
            }
            const float depth = dot(target - pos, fwd);
            const f3 shift = ((px0 + px1) * 0.5f * depth) * r + ((py0 + py1) * 0.5f * depth) * u;
            pos = pos + shift;
            target = target + shift;
            float dist = 0.f;
            for (int i = 0; i < 8; i++) {
                const f3 v = corner(i) - target;
                const float z = dot(v, fwd);
                dist = std::max(dist, std::max(std::fabs(dot(v, r)) / tan_hx - z, std::fabs(dot(v, u)) / tan_hy - z));
            }
            pos = target - fwd * dist;
        }
        const f3 fwd = normalize(target - pos);
        for (int k = 0; k < 3; k++) {
            c.eye[k] = (&pos.x)[k];
            c.forward[k] = (&fwd.x)[k];
            c.up[k] = (&up.x)[k];
        }
        c.vfov = vfov;
        c.aspect = aspect;
        c.has_resolution = false;
        c.present = true;
    }
};

} // namespace

mrt_scene_desc Scene::desc() const {
    mrt_scene_desc d{};
    d.vertices = vertices.data();
    d.vertex_count = uint32_t(vertices.size());
    d.indices = indices.data();
    d.index_count = uint32_t(indices.size());
    d.geometries = geometries.data();
    d.geometry_count = uint32_t(geometries.size());
    d.materials = materials.data();
    d.material_count = uint32_t(materials.size());
    if (!env_rgba.empty()) {
        d.env.type = MRT_ENV_EQUAL_AREA_OCT;
        d.env.wrap = MRT_WRAP_CLAMP;
        d.env.width = env_width;
        d.env.height = env_height;
        d.env.rgba = env_rgba.data();
        d.env.intensity = env_scale;
        std::memcpy(d.env.to_local, env_to_local, sizeof(env_to_local));
    } else {
        d.env.type = MRT_ENV_NONE;
    }
    return d;
}

Scene load(const std::string& path) {
    Scene sc;
    Loader l(sc);
    l.parse_file(path);
    l.finish();
    return sc;
}

bool read_pfm(const std::string& path, std::vector<float>& rgba, uint32_t& width, uint32_t& height) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    std::string magic;
    double scale;
    in >> magic >> width >> height >> scale;
    in.get();
    if (!in || (magic != "PF" && magic != "Pf") || width == 0 || height == 0) return false;
    const int ch = magic == "PF" ? 3 : 1;
    std::vector<float> raw(size_t(width) * height * ch);
    in.read(reinterpret_cast<char*>(raw.data()), std::streamsize(raw.size() * 4));
    if (!in) return false;
    const bool big = scale > 0; // PFM: positive scale = big endian
    if (big)
        for (float& f : raw) {
            uint32_t u;
            std::memcpy(&u, &f, 4);
            u = __builtin_bswap32(u);
            std::memcpy(&f, &u, 4);
        }
    rgba.assign(size_t(width) * height * 4, 1.f);
    for (uint32_t y = 0; y < height; y++) // PFM rows run bottom to top
        for (uint32_t x = 0; x < width; x++)
            for (int c = 0; c < 3; c++)
                rgba[(size_t(y) * width + x) * 4 + c] = raw[(size_t(height - 1 - y) * width + x) * ch + (ch == 3 ? c : 0)];
    return true;
}

bool write_pfm(const std::string& path, const float* rgba, uint32_t width, uint32_t height) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    std::fprintf(f, "PF\n%u %u\n-1.0\n", width, height);
    std::vector<float> row(size_t(width) * 3);
    bool ok = true;
    for (uint32_t y = 0; y < height && ok; y++) {
        const float* src = rgba + size_t(height - 1 - y) * width * 4;
        for (uint32_t x = 0; x < width; x++)
            for (int c = 0; c < 3; c++) row[3 * x + c] = src[4 * x + c];
        ok = std::fwrite(row.data(), 4, row.size(), f) == row.size();
    }
    return std::fclose(f) == 0 && ok;
}

} // namespace mrt::pbrt
