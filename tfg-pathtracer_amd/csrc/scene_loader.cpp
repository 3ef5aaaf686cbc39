// Scene ingestion with the reference loader's semantics, written from scratch:
//   loadScene            SceneLoader.hpp:7-143
//   .mtl parsing         ObjLoader.hpp:33-69, 191-205
//   .obj parsing         ObjLoader.hpp:71-189  (z negated, normals normalised)
//   Vector3(std::string) Vector.hpp:14-60      (first three numbers of a text)
//   Texture(path, cs)    Texture.hpp:41-68     (vertical flip, LDR gamma via fastPow)
//   HDRI(path|color)     HDRI.hpp:22-60        (unflipped, CDF at construction)
// Deliberate supersets, each harmless for files the reference accepts:
//   '/' and '\\' both separate paths; relative map paths are tried against the
//   working directory (as the reference) and then the scene directory;
//   numbers may use exponents; faces with more than 3 vertices are fanned;
//   negative (relative) OBJ indices are resolved; MTL keys `Pr` / `Pm` set
//   scalar roughness / metallic (the reference has no scalar keys for them).
// Tangents: Mikkelsen's tangent space per mesh object (tangent_space.cpp),
// the reference's MikkTSpace convention (ObjLoader.hpp:167-168,
// mikktspaceCallback.hpp:27-99); they feed normal maps and the
// TANGENT/BITANGENT passes.
#include "scene.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

namespace eleven {

// ------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------
static bool fileExists(const std::string& p) {
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

static std::string normSep(std::string p) {
    for (auto& c : p)
        if (c == '\\') c = '/';
    return p;
}

static std::string joinPath(const std::string& dir, const std::string& name) {
    if (dir.empty()) return name;
    if (dir.back() == '/') return dir + name;
    return dir + "/" + name;
}

static std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

static bool readWholeFile(const std::string& path, std::string& out) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    std::ostringstream ss;
    ss << in.rdbuf();
    out = ss.str();
    return true;
}

// Vector.hpp:14-60: the first three numbers found in `s`, missing ones are 0.
static Vector3 vec3FromText(const std::string& s) {
    float v[3] = {0, 0, 0};
    int n = 0;
    size_t i = 0;
    while (i < s.size() && n < 3) {
        char c = s[i];
        if (std::isdigit((unsigned char)c) || c == '.' || c == '-' || c == '+') {
            const char* start = s.c_str() + i;
            char* end = nullptr;
            float f = std::strtof(start, &end);
            if (end == start) {  // a lone '-' or '.', skip it
                i++;
                continue;
            }
            v[n++] = f;
            i = (size_t)(end - s.c_str());
        } else {
            i++;
        }
    }
    return Vector3(v[0], v[1], v[2]);
}

// Vector.hpp:173-176 normalized(): zero vectors are returned unchanged.
static Vector3 normalizedRef(const Vector3& a) {
    float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    if (l == 0) return a;
    return Vector3(a.x / l, a.y / l, a.z / l);
}

// ------------------------------------------------------------------------
// fastPow (Ankerl 2012)
// ------------------------------------------------------------------------
double fastPow(double a, double b) {
    uint64_t bits;
    std::memcpy(&bits, &a, 8);
    int32_t hi = (int32_t)(bits >> 32);
    int32_t nhi = (int32_t)(b * (double)(hi - 1072632447) + 1072632447);
    bits = ((uint64_t)(uint32_t)nhi) << 32;
    double r;
    std::memcpy(&r, &bits, 8);
    return r;
}

// ------------------------------------------------------------------------
// minimal JSON (objects, arrays, strings, numbers, literals) keeping the raw
// text of every value so that Vector3-from-text semantics can be applied.
// ------------------------------------------------------------------------
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    std::string raw;
    const JVal* get(const std::string& k) const {
        for (auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    double asNum() const {
        if (kind == NUM) return num;
        if (kind == BOOL) return b ? 1.0 : 0.0;
        if (kind == STR) return std::strtod(str.c_str(), nullptr);
        return 0;
    }
    bool asBool() const {
        if (kind == BOOL) return b;
        if (kind == NUM) return num != 0;
        if (kind == STR) return str == "true";
        return false;
    }
};

struct JParser {
    const std::string& s;
    size_t p = 0;
    explicit JParser(const std::string& src) : s(src) {}
    void ws() {
        while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
    }
    [[noreturn]] void fail(const char* what) {
        throw LoadError{std::string("scene.json: ") + what + " at offset " + std::to_string(p)};
    }
    std::string parseString() {
        if (s[p] != '"') fail("expected string");
        p++;
        std::string out;
        while (p < s.size() && s[p] != '"') {
            if (s[p] == '\\' && p + 1 < s.size()) {
                p++;
                char e = s[p];
                out.push_back(e == 'n' ? '\n' : e == 't' ? '\t' : e);
            } else {
                out.push_back(s[p]);
            }
            p++;
        }
        if (p >= s.size()) fail("unterminated string");
        p++;
        return out;
    }
    JVal parse() {
        ws();
        if (p >= s.size()) fail("unexpected end");
        size_t start = p;
        JVal v;
        char c = s[p];
        if (c == '{') {
            v.kind = JVal::OBJ;
            p++;
            ws();
            if (p < s.size() && s[p] == '}') {
                p++;
            } else {
                while (true) {
                    ws();
                    std::string k = parseString();
                    ws();
                    if (p >= s.size() || s[p] != ':') fail("expected ':'");
                    p++;
                    JVal child = parse();
                    v.obj.emplace_back(k, std::move(child));
                    ws();
                    if (p < s.size() && s[p] == ',') {
                        p++;
                        continue;
                    }
                    if (p < s.size() && s[p] == '}') {
                        p++;
                        break;
                    }
                    fail("expected ',' or '}'");
                }
            }
        } else if (c == '[') {
            v.kind = JVal::ARR;
            p++;
            ws();
            if (p < s.size() && s[p] == ']') {
                p++;
            } else {
                while (true) {
                    v.arr.push_back(parse());
                    ws();
                    if (p < s.size() && s[p] == ',') {
                        p++;
                        continue;
                    }
                    if (p < s.size() && s[p] == ']') {
                        p++;
                        break;
                    }
                    fail("expected ',' or ']'");
                }
            }
        } else if (c == '"') {
            v.kind = JVal::STR;
            v.str = parseString();
        } else if (s.compare(p, 4, "true") == 0) {
            v.kind = JVal::BOOL;
            v.b = true;
            p += 4;
        } else if (s.compare(p, 5, "false") == 0) {
            v.kind = JVal::BOOL;
            v.b = false;
            p += 5;
        } else if (s.compare(p, 4, "null") == 0) {
            v.kind = JVal::NUL;
            p += 4;
        } else {
            const char* b = s.c_str() + p;
            char* e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) fail("bad value");
            v.kind = JVal::NUM;
            p += (size_t)(e - b);
        }
        v.raw = s.substr(start, p - start);
        return v;
    }
};

// ------------------------------------------------------------------------
// images
// ------------------------------------------------------------------------
static uint32_t rd32(const unsigned char* b) {
    return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
}
static uint16_t rd16(const unsigned char* b) { return (uint16_t)(b[0] | (b[1] << 8)); }

// Uncompressed 24/32-bit BMP -> 8-bit RGB rows, bottom row first (the order
// stb returns with vertical flip on, which the reference's Texture sets).
static bool readBMP(const std::string& path, int& w, int& h, std::vector<unsigned char>& rgb,
                    std::string& err) {
    std::string buf;
    if (!readWholeFile(path, buf)) {
        err = "cannot open " + path;
        return false;
    }
    const unsigned char* d = (const unsigned char*)buf.data();
    if (buf.size() < 54 || d[0] != 'B' || d[1] != 'M') {
        err = "not a BMP file: " + path;
        return false;
    }
    uint32_t off = rd32(d + 10);
    int32_t bw = (int32_t)rd32(d + 18);
    int32_t bh = (int32_t)rd32(d + 22);
    uint16_t bpp = rd16(d + 28);
    uint32_t comp = rd32(d + 30);
    if ((bpp != 24 && bpp != 32) || (comp != 0 && comp != 3) || bw <= 0 || bh == 0) {
        err = "unsupported BMP (need uncompressed 24/32-bit): " + path;
        return false;
    }
    bool topDown = bh < 0;
    h = topDown ? -bh : bh;
    w = bw;
    size_t stride = ((size_t)w * (bpp / 8) + 3) & ~(size_t)3;
    if ((size_t)off + stride * (size_t)h > buf.size()) {
        err = "truncated BMP: " + path;
        return false;
    }
    rgb.assign((size_t)w * h * 3, 0);
    for (int r = 0; r < h; r++) {
        // output row r = r-th row from the bottom
        int fileRow = topDown ? (h - 1 - r) : r;
        const unsigned char* src = d + off + stride * (size_t)fileRow;
        unsigned char* dst = rgb.data() + (size_t)r * w * 3;
        for (int x = 0; x < w; x++) {
            const unsigned char* px = src + (size_t)x * (bpp / 8);
            dst[3 * x + 0] = px[2];
            dst[3 * x + 1] = px[1];
            dst[3 * x + 2] = px[0];
        }
    }
    return true;
}

// Radiance RGBE (.hdr) with new-style RLE scanlines; rows in file order
// (top first), which is what the reference gets for its HDRI.
static bool readHDR(const std::string& path, int& w, int& h, std::vector<float>& rgb,
                    std::string& err) {
    std::string buf;
    if (!readWholeFile(path, buf)) {
        err = "cannot open " + path;
        return false;
    }
    size_t p = 0;
    auto line = [&](std::string& out) -> bool {
        if (p >= buf.size()) return false;
        size_t e = buf.find('\n', p);
        if (e == std::string::npos) e = buf.size();
        out = buf.substr(p, e - p);
        p = e + 1;
        return true;
    };
    std::string l;
    if (!line(l) || (l.rfind("#?RADIANCE", 0) != 0 && l.rfind("#?RGBE", 0) != 0)) {
        err = "not a Radiance HDR file: " + path;
        return false;
    }
    bool fmtOk = false;
    while (line(l)) {
        if (l.empty()) break;
        if (l == "FORMAT=32-bit_rle_rgbe") fmtOk = true;
    }
    if (!fmtOk) {
        err = "unsupported HDR format: " + path;
        return false;
    }
    if (!line(l)) {
        err = "missing HDR resolution: " + path;
        return false;
    }
    char ya[4] = {0}, xa[4] = {0};
    if (std::sscanf(l.c_str(), "%3s %d %3s %d", ya, &h, xa, &w) != 4 || std::strcmp(ya, "-Y") != 0 ||
        std::strcmp(xa, "+X") != 0 || w <= 0 || h <= 0) {
        err = "unsupported HDR orientation (need -Y H +X W): " + path;
        return false;
    }
    rgb.assign((size_t)w * h * 3, 0.f);
    std::vector<unsigned char> scan((size_t)w * 4);
    const unsigned char* d = (const unsigned char*)buf.data();
    size_t n = buf.size();
    for (int y = 0; y < h; y++) {
        bool rle = w >= 8 && w < 32768 && p + 4 <= n && d[p] == 2 && d[p + 1] == 2 &&
                   (((int)d[p + 2] << 8) | d[p + 3]) == w && !(d[p + 2] & 0x80);
        if (rle) {
            p += 4;
            for (int c = 0; c < 4; c++) {
                int x = 0;
                while (x < w) {
                    if (p >= n) {
                        err = "truncated HDR: " + path;
                        return false;
                    }
                    int count = d[p++];
                    if (count > 128) {
                        count -= 128;
                        if (p >= n || x + count > w) {
                            err = "corrupt HDR run: " + path;
                            return false;
                        }
                        unsigned char v = d[p++];
                        for (int k = 0; k < count; k++) scan[(size_t)(x++) * 4 + c] = v;
                    } else {
                        if (count == 0 || p + count > n || x + count > w) {
                            err = "corrupt HDR run: " + path;
                            return false;
                        }
                        for (int k = 0; k < count; k++) scan[(size_t)(x++) * 4 + c] = d[p++];
                    }
                }
            }
        } else {
            if (p + (size_t)w * 4 > n) {
                err = "truncated HDR: " + path;
                return false;
            }
            std::memcpy(scan.data(), d + p, (size_t)w * 4);
            p += (size_t)w * 4;
        }
        for (int x = 0; x < w; x++) {
            const unsigned char* e = &scan[(size_t)x * 4];
            float* o = &rgb[((size_t)y * w + x) * 3];
            if (e[3] == 0) {
                o[0] = o[1] = o[2] = 0.f;
            } else {
                float f = std::ldexp(1.0f, (int)e[3] - (128 + 8));
                o[0] = e[0] * f;
                o[1] = e[1] * f;
                o[2] = e[2] * f;
            }
        }
    }
    return true;
}

Texture loadTexture(const std::string& path, CS colorSpace) {
    Texture t;
    t.path = path;
    std::string p = normSep(path);
    std::string lower = p;
    std::transform(lower.begin(), lower.end(), lower.begin(), ::tolower);
    std::string err;
    if (lower.size() >= 4 && lower.compare(lower.size() - 4, 4, ".hdr") == 0) {
        // stb returns HDR data linear, unaffected by the LDR gamma.
        std::vector<float> rgb;
        int w, h;
        if (!readHDR(p, w, h, rgb, err)) throw LoadError{err};
        // stb flips on load for textures; rows of readHDR are top first.
        t.width = w;
        t.height = h;
        t.data.resize(rgb.size());
        for (int y = 0; y < h; y++)
            std::memcpy(&t.data[(size_t)y * w * 3], &rgb[(size_t)(h - 1 - y) * w * 3], sizeof(float) * w * 3);
        return t;
    }
    std::vector<unsigned char> rgb;
    int w, h;
    if (!readBMP(p, w, h, rgb, err)) throw LoadError{err};
    t.width = w;
    t.height = h;
    t.data.resize(rgb.size());
    float lut[256];
    ldrTable(colorSpace, lut);
    for (size_t i = 0; i < rgb.size(); i++) t.data[i] = lut[rgb[i]];
    return t;
}

// stb_image.h:1855-1870 (modified): fastPow(v/255.0f, gamma) * scale.
void ldrTable(CS colorSpace, float out[256]) {
    const double gamma = colorSpace == sRGB ? (double)2.2f : 1.0;
    for (int i = 0; i < 256; i++) out[i] = (float)(fastPow((double)(i / 255.0f), gamma) * 1.0f);
}

Vector3 texelAt(const Texture& t, int x, int y) {
    x = (int)(t.xTile * (x + t.xOffset * t.width)) % t.width;
    y = (int)(t.yTile * (y + t.yOffset * t.height)) % t.height;
    size_t i = 3 * ((size_t)y * t.width + x);
    return Vector3(t.data[i + 0], t.data[i + 1], t.data[i + 2]);
}

void generateCDF(HDRI& h) {
    const Texture& tx = h.texture;
    size_t n = (size_t)tx.width * tx.height;
    h.cdf.assign(n + 1, 0.f);
    float sum = 0;
    for (int j = 0; j < tx.height; j++)
        for (int i = 0; i < tx.width; i++) {
            Vector3 d = texelAt(tx, i, j);
            sum += d.x + d.y + d.z;
        }
    h.radianceSum = sum;
    size_t c = 0;
    for (int j = 0; j < tx.height; j++)
        for (int i = 0; i < tx.width; i++) {
            Vector3 d = texelAt(tx, i, j);
            h.cdf[c + 1] = h.cdf[c] + (d.x + d.y + d.z) / sum;
            c++;
        }
}

HDRI makeConstantHDRI(Vector3 color) {
    HDRI h;
    h.texture.width = 1024;
    h.texture.height = 1024;
    h.texture.data.resize((size_t)1024 * 1024 * 3);
    for (size_t i = 0; i < (size_t)1024 * 1024; i++) {
        h.texture.data[3 * i + 0] = color.x;
        h.texture.data[3 * i + 1] = color.y;
        h.texture.data[3 * i + 2] = color.z;
    }
    generateCDF(h);
    return h;
}

HDRI loadHDRIFile(const std::string& path) {
    HDRI h;
    std::string err;
    std::vector<float> rgb;
    int w, hh;
    if (!readHDR(normSep(path), w, hh, rgb, err)) throw LoadError{err};
    h.texture.width = w;
    h.texture.height = hh;
    h.texture.path = path;
    h.texture.data = std::move(rgb);
    generateCDF(h);
    return h;
}

// ------------------------------------------------------------------------
// MTL / OBJ
// ------------------------------------------------------------------------
struct UnloadedMaterial {
    Material mat;
    std::map<std::string, std::string> maps;
};

static std::string secondWord(const std::string& line) {   // ObjLoader.hpp:28-31
    std::string::size_type in = line.find_first_of(' ');
    if (in == std::string::npos) return "";
    return trim(line.substr(in + 1));
}

static std::vector<UnloadedMaterial> loadMtls(const std::string& path) {
    std::vector<UnloadedMaterial> out;
    std::ifstream in(path);
    if (!in) throw LoadError{"cannot open " + path};
    std::string raw;
    UnloadedMaterial* cur = nullptr;
    static const char* mapNames[] = {"map_Kd", "map_Ns", "map_Bump", "refl"};
    while (std::getline(in, raw)) {
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        if (raw.find("newmtl") != std::string::npos) {
            out.emplace_back();
            cur = &out.back();
            cur->mat.name = secondWord(raw);
            continue;
        }
        if (!cur) continue;
        std::string line = raw;
        char c0 = line.size() > 0 ? line[0] : '\0';
        char c1 = line.size() > 1 ? line[1] : '\0';
        if (c0 == 'K' && c1 == 'd') cur->mat.albedo = vec3FromText(line.substr(2));
        if (c0 == 'K' && c1 == 's') cur->mat.specular = vec3FromText(line.substr(2)).x;
        if (c0 == 'K' && c1 == 'e') cur->mat.emission = vec3FromText(line.substr(2));
        if (c0 == 'N' && c1 == 'i') cur->mat.eta = vec3FromText(secondWord(line)).x;
        if (c0 == 'd') cur->mat.opacity = Vector3(vec3FromText(secondWord(line)).x);
        if (c0 == 'P' && c1 == 'r') cur->mat.roughness = vec3FromText(line.substr(2)).x;
        if (c0 == 'P' && c1 == 'm') cur->mat.metallic = vec3FromText(line.substr(2)).x;
        for (const char* m : mapNames)
            if (line.find(m) != std::string::npos) cur->maps[m] = secondWord(line);
    }
    return out;
}


static long parseIndex(const char* s, size_t n, long count) {
    // 1-based; negative = relative to the end (superset of the reference).
    std::string tok(s, n);
    if (tok.empty()) return -1;
    long v = std::strtol(tok.c_str(), nullptr, 10);
    if (v < 0) v = count + v + 1;
    return v - 1;
}

static void loadObjs(const std::string& path, std::vector<MeshObject>& meshes,
                     std::vector<std::vector<Tri>>& meshTris) {
    std::ifstream in(path);
    if (!in) throw LoadError{"cannot open " + path};
    std::vector<Vector3> V, VT, VN;
    std::string line;
    MeshObject* cur = nullptr;
    std::vector<Tri>* curTris = nullptr;
    long lineNo = 0;
    while (std::getline(in, line)) {
        lineNo++;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) continue;
        if (line[0] == 'o') {   // ObjLoader.hpp:179-186: each 'o' starts an object
            meshes.emplace_back();
            meshTris.emplace_back();
            cur = &meshes.back();
            curTris = &meshTris.back();
            cur->name = secondWord(line);
            continue;
        }
        if (!cur) continue;     // the reference ignores everything before the first 'o'
        if (line.find("usemtl") != std::string::npos) cur->matName = secondWord(line);
        if (line[0] == 'v' && line.size() > 1 && line[1] == ' ') {
            Vector3 v = vec3FromText(line.substr(2));
            V.push_back(Vector3(v.x, v.y, -v.z));
        } else if (line[0] == 'v' && line.size() > 1 && line[1] == 't') {
            VT.push_back(vec3FromText(line.substr(2)));
        } else if (line[0] == 'v' && line.size() > 1 && line[1] == 'n') {
            Vector3 n = vec3FromText(line.substr(2));
            VN.push_back(Vector3(n.x, n.y, -n.z));
        } else if (line[0] == 'f') {
            // vertex groups: runs of [0-9/-] separated by anything else
            std::vector<long> vi, ti, ni;
            size_t i = 1;
            while (i < line.size()) {
                while (i < line.size() && !(std::isdigit((unsigned char)line[i]) || line[i] == '/' || line[i] == '-')) i++;
                if (i >= line.size()) break;
                size_t s = i;
                while (i < line.size() && (std::isdigit((unsigned char)line[i]) || line[i] == '/' || line[i] == '-')) i++;
                const char* g = line.c_str() + s;
                size_t gl = i - s;
                size_t a = 0;
                long idx[3] = {-1, -1, -1};
                for (int k = 0; k < 3 && a <= gl; k++) {
                    size_t b = a;
                    while (b < gl && g[b] != '/') b++;
                    long cnt = k == 0 ? (long)V.size() : k == 1 ? (long)VT.size() : (long)VN.size();
                    idx[k] = parseIndex(g + a, b - a, cnt);
                    a = b + 1;
                    if (b >= gl) break;
                }
                vi.push_back(idx[0]);
                ti.push_back(idx[1]);
                ni.push_back(idx[2]);
            }
            if (vi.size() < 3)
                throw LoadError{path + ":" + std::to_string(lineNo) + ": face with fewer than 3 vertices"};
            for (size_t k = 1; k + 1 < vi.size(); k++) {
                Tri tri;
                size_t c[3] = {0, k, k + 1};
                for (int j = 0; j < 3; j++) {
                    long a = vi[c[j]], b = ti[c[j]], n = ni[c[j]];
                    if (a < 0 || a >= (long)V.size())
                        throw LoadError{path + ":" + std::to_string(lineNo) + ": vertex index out of range"};
                    tri.vertices[j] = V[a];
                    if (b >= 0) {
                        if (b >= (long)VT.size())
                            throw LoadError{path + ":" + std::to_string(lineNo) + ": uv index out of range"};
                        tri.uv[j] = VT[b];
                    }
                    if (n >= 0) {
                        if (n >= (long)VN.size())
                            throw LoadError{path + ":" + std::to_string(lineNo) + ": normal index out of range"};
                        tri.normals[j] = normalizedRef(VN[n]);
                    }
                }
                curTris->push_back(tri);
            }
        }
    }
    for (auto& t : meshTris) mikkTangents(t);   // per mesh object, as ObjLoader.hpp:167-168
}

void Scene::addMeshObject(MeshObject mo, const std::vector<Tri>& meshTris) {
    mo.objectID = meshObjectCount();
    mo.triFirst = (unsigned int)tris.size();
    mo.triCount = (unsigned int)meshTris.size();
    for (const Tri& t0 : meshTris) {
        Tri t = t0;
        t.objectID = mo.objectID;
        tris.push_back(t);
    }
    meshObjects.push_back(mo);
}

static std::string resolveMapPath(const std::string& dir, const std::string& raw) {
    std::string p = normSep(raw);
    if (fileExists(p)) return p;
    std::string q = joinPath(dir, p);
    if (fileExists(q)) return q;
    return p;  // loadTexture reports the failure
}

Scene loadScene(const std::string& dirIn) {
    std::string dir = normSep(dirIn);
    while (dir.size() > 1 && dir.back() == '/') dir.pop_back();
    Scene scene;

    std::string js;
    if (!readWholeFile(joinPath(dir, "scene.json"), js)) throw LoadError{"cannot open " + joinPath(dir, "scene.json")};
    JParser jp(js);
    JVal root = jp.parse();
    if (root.kind != JVal::OBJ) throw LoadError{"scene.json: top level is not an object"};

    // Camera (SceneLoader.hpp:24-46)
    const JVal* cam = root.get("camera");
    if (!cam || cam->kind != JVal::OBJ) throw LoadError{"scene.json: missing camera"};
    auto num = [&](const JVal* o, const char* k, double def) {
        const JVal* v = o ? o->get(k) : nullptr;
        return v ? v->asNum() : def;
    };
    int xRes = (int)num(cam, "xRes", 1280), yRes = (int)num(cam, "yRes", 720);
    if (xRes <= 0 || yRes <= 0) throw LoadError{"scene.json: camera resolution must be positive"};
    scene.camera = Camera((unsigned)xRes, (unsigned)yRes);
    scene.camera.focalLength = (float)num(cam, "focalLength", 0.035);
    scene.camera.focusDistance = (float)num(cam, "focusDistance", 1000000);
    scene.camera.aperture = (float)num(cam, "aperture", 2.8);
    const JVal* bk = cam->get("bokeh");
    scene.camera.bokeh = bk ? bk->asBool() : false;
    const JVal* cp = cam->get("position");
    const JVal* cr = cam->get("rotation");
    scene.camera.position = Vector3((float)num(cp, "x", 0), (float)num(cp, "y", 0), (float)num(cp, "z", 0));
    scene.camera.rotation = Vector3((float)num(cr, "x", 0), (float)num(cr, "y", 0), (float)num(cr, "z", 0));

    // HDRI (SceneLoader.hpp:51-65).  The CDF is built by the constructor,
    // before the offsets are applied, and "yOffset" writes xOffset exactly as
    // SceneLoader.hpp:63-64 does.
    const JVal* hd = root.get("hdri");
    if (hd && hd->get("name")) {
        const JVal* nm = hd->get("name");
        scene.hdri = loadHDRIFile(joinPath(joinPath(dir, "HDRI"), nm->str + ".hdr"));
    } else if (hd && hd->get("color")) {
        const JVal* c = hd->get("color");
        scene.hdri = makeConstantHDRI(Vector3((float)num(c, "r", 0), (float)num(c, "g", 0), (float)num(c, "b", 0)));
    } else {
        // HDRI() default: a 1x1 black texture (HDRI.hpp:18-20, Texture.hpp:64-70).
        scene.hdri.texture.width = 1;
        scene.hdri.texture.height = 1;
        scene.hdri.texture.data.assign(3, 0.f);
        generateCDF(scene.hdri);
    }
    if (hd && hd->get("xOffset")) scene.hdri.texture.xOffset = (float)hd->get("xOffset")->asNum();
    if (hd && hd->get("yOffset")) scene.hdri.texture.xOffset = (float)hd->get("yOffset")->asNum();

    // Materials (SceneLoader.hpp:69-107)
    std::string mtlPath = joinPath(dir, "scene.mtl");
    std::vector<UnloadedMaterial> umtls;
    if (fileExists(mtlPath)) umtls = loadMtls(mtlPath);
    for (auto& um : umtls) {
        for (auto& kv : um.maps) {
            const std::string& key = kv.first;
            std::string mapPath = resolveMapPath(dir, kv.second);
            CS cs = LINEAR;
            int textureId = scene.textureCount();
            bool dup = false;
            for (int j = 0; j < scene.textureCount(); j++)
                if (scene.textures[j].path == mapPath) {
                    textureId = j;
                    dup = true;
                }
            if (key == "map_Kd") {
                um.mat.albedoTextureID = textureId;
                cs = sRGB;
            }
            if (key == "map_Ns") um.mat.roughnessTextureID = textureId;
            if (key == "refl") um.mat.metallicTextureID = textureId;
            if (key == "map_Bump") um.mat.normalTextureID = textureId;
            if (!dup) scene.textures.push_back(loadTexture(mapPath, cs));
        }
        scene.materials.push_back(um.mat);
    }

    // Objects (SceneLoader.hpp:112-123)
    std::vector<MeshObject> objs;
    std::vector<std::vector<Tri>> objTris;
    loadObjs(joinPath(dir, "scene.obj"), objs, objTris);
    for (size_t i = 0; i < objs.size(); i++) {
        objs[i].materialID = 0;
        for (int j = 0; j < scene.materialCount(); j++)
            if (scene.materials[j].name == objs[i].matName) objs[i].materialID = j;
        scene.addMeshObject(objs[i], objTris[i]);
    }
    if (scene.materials.empty()) scene.materials.push_back(Material());  // materialID 0 must exist

    // Point lights (SceneLoader.hpp:127-141)
    const JVal* pl = root.get("pointLights");
    if (pl && pl->kind == JVal::ARR) {
        for (const JVal& l : pl->arr) {
            const JVal* p = l.get("position");
            const JVal* r = l.get("radiance");
            PointLight L;
            L.position = p ? vec3FromText(p->raw) : Vector3();
            L.radiance = r ? vec3FromText(r->raw) : Vector3();
            scene.pointLights.push_back(L);
        }
    }
    return scene;
}

// ------------------------------------------------------------------------
// writers
// ------------------------------------------------------------------------
static uint32_t crcTable[256];
static bool crcInit = false;
static uint32_t crc32(const unsigned char* d, size_t n, uint32_t c = 0xffffffffu) {
    if (!crcInit) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t v = i;
            for (int k = 0; k < 8; k++) v = (v & 1) ? 0xedb88320u ^ (v >> 1) : v >> 1;
            crcTable[i] = v;
        }
        crcInit = true;
    }
    for (size_t i = 0; i < n; i++) c = crcTable[(c ^ d[i]) & 0xff] ^ (c >> 8);
    return c;
}

bool writePNG(const std::string& path, int w, int h, const unsigned char* rgba) {
    // zlib stream of stored (uncompressed) deflate blocks: valid PNG, no zlib needed.
    std::vector<unsigned char> raw;
    raw.reserve((size_t)h * (w * 4 + 1));
    for (int y = 0; y < h; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba + (size_t)y * w * 4, rgba + (size_t)(y + 1) * w * 4);
    }
    std::vector<unsigned char> z = {0x78, 0x01};
    size_t pos = 0;
    uint32_t a = 1, b = 0;
    for (unsigned char c : raw) {
        a = (a + c) % 65521;
        b = (b + a) % 65521;
    }
    do {
        size_t len = std::min<size_t>(65535, raw.size() - pos);
        bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back(len & 0xff);
        z.push_back((len >> 8) & 0xff);
        z.push_back(~len & 0xff);
        z.push_back((~len >> 8) & 0xff);
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t ad = (b << 16) | a;
    for (int s = 24; s >= 0; s -= 8) z.push_back((ad >> s) & 0xff);

    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::fwrite(sig, 1, 8, f);
    auto chunk = [&](const char* type, const std::vector<unsigned char>& data) {
        unsigned char len[4] = {(unsigned char)(data.size() >> 24), (unsigned char)(data.size() >> 16),
                                (unsigned char)(data.size() >> 8), (unsigned char)data.size()};
        std::fwrite(len, 1, 4, f);
        std::vector<unsigned char> td(type, type + 4);
        td.insert(td.end(), data.begin(), data.end());
        std::fwrite(td.data(), 1, td.size(), f);
        uint32_t c = crc32(td.data(), td.size()) ^ 0xffffffffu;
        unsigned char cb[4] = {(unsigned char)(c >> 24), (unsigned char)(c >> 16), (unsigned char)(c >> 8),
                               (unsigned char)c};
        std::fwrite(cb, 1, 4, f);
    };
    std::vector<unsigned char> ihdr = {(unsigned char)(w >> 24), (unsigned char)(w >> 16), (unsigned char)(w >> 8),
                                       (unsigned char)w, (unsigned char)(h >> 24), (unsigned char)(h >> 16),
                                       (unsigned char)(h >> 8), (unsigned char)h, 8, 6, 0, 0, 0};
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", {});
    bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok;
}

bool writeBMP(const std::string& path, int w, int h, const unsigned char* rgba) {
    size_t stride = ((size_t)w * 3 + 3) & ~(size_t)3;
    size_t size = 54 + stride * h;
    std::vector<unsigned char> b(size, 0);
    auto w32 = [&](size_t o, uint32_t v) {
        b[o] = v & 0xff;
        b[o + 1] = (v >> 8) & 0xff;
        b[o + 2] = (v >> 16) & 0xff;
        b[o + 3] = (v >> 24) & 0xff;
    };
    b[0] = 'B';
    b[1] = 'M';
    w32(2, (uint32_t)size);
    w32(10, 54);
    w32(14, 40);
    w32(18, (uint32_t)w);
    w32(22, (uint32_t)h);
    b[26] = 1;
    b[28] = 24;
    w32(34, (uint32_t)(stride * h));
    for (int y = 0; y < h; y++) {
        const unsigned char* src = rgba + (size_t)(h - 1 - y) * w * 4;  // bottom row first
        unsigned char* dst = b.data() + 54 + stride * y;
        for (int x = 0; x < w; x++) {
            dst[3 * x + 0] = src[4 * x + 2];
            dst[3 * x + 1] = src[4 * x + 1];
            dst[3 * x + 2] = src[4 * x + 0];
        }
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
    std::fclose(f);
    return ok;
}

}  // namespace eleven
