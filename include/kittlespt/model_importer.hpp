// kittlespt/model_importer.hpp — glTF 2.0 files (.gltf / .glb) into a
// KittlesPT::BasicScene for C++ applications, with the conventions of the
// reference's SampleApp importer (SampleApp/src/sample_app/model_importer.hpp:
// 51-456, ModelImporter::loadGLTFfromFile), and the same scene as this
// library's Python importer (kittlespt_amd/gltf.py) for the same file.
//
// What is read, as the reference reads it:
//  * the first scene's top-level nodes (:83-100; children are not visited,
//    parseNode returns at once, :421-443); a node with a camera becomes a
//    CameraSceneEntity whose view is the inverse of T * R(normalised
//    quaternion) or of `matrix`, with the camera's yfov (:380-419); a node
//    with a mesh becomes one MeshSceneEntity whose model matrix is `matrix`,
//    else T * R * S (:155-198, glm's quaternion-to-matrix form in float);
//  * every primitive's triangles expanded through its indices: POSITION,
//    NORMAL and TEXCOORD_0 per corner, the primitive's material per triangle
//    (:225-274);
//  * images as RGB8 textures (stb_image with 3 channels, :276-318) from
//    buffer views or data URIs, or files next to the .gltf (the reference
//    looks in a fixed "../models/" directory instead);
//  * materials: base colour factor (RGB) and texture, metallic-roughness
//    factors and texture, normal texture and scale, emissive factor and
//    texture, KHR_materials_transmission, KHR_materials_ior (default 1.45)
//    and KHR_materials_emissive_strength (:320-377).
// Defined where the reference is not: accessors are read as the glTF
// specification lays them out (accessor and buffer-view offsets, byte
// stride, 8/16/32-bit indices; the reference assumes tightly packed
// attributes and 16-bit indices), a primitive without indices is read in
// vertex order and one without a material gets material 0, a file without
// materials gets one default material, PNG is the only image format decoded
// (kpt_png_decode_rgb8), and the camera view is inverted in double precision
// and rounded to float.  A malformed file makes loadGLTFfromFile return false
// with the reason in lastError(); the scene is then left unchanged.
//
// Header-only; link the application against libkpt_hip.so (PNG decoding).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../kpt_abi.h"
#include "pod_types.hpp"

namespace KittlesPT {

namespace importer_detail {

// ---------------------------------------------------------------- JSON (RFC 8259)
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    const Json* get(const char* key) const {
        if (kind != Object) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    const Json* at(size_t i) const { return (kind == Array && i < arr.size()) ? &arr[i] : nullptr; }
    size_t size() const { return kind == Array ? arr.size() : 0; }
    double number(const char* key, double def) const {
        const Json* v = get(key);
        return (v && v->kind == Number) ? v->num : def;
    }
    long long integer(const char* key, long long def) const {
        const Json* v = get(key);
        // out-of-range values (beyond what any index or count can be) read as the default
        return (v && v->kind == Number && std::fabs(v->num) < 9.0e18) ? (long long)v->num : def;
    }
    std::string string(const char* key) const {
        const Json* v = get(key);
        return (v && v->kind == String) ? v->str : std::string();
    }
};

class JsonParser {
public:
    JsonParser(const char* p, size_t n) : p_(p), end_(p + n) {}
    Json parse() {
        Json v = value(0);
        ws();
        if (p_ != end_) fail("trailing characters");
        return v;
    }

private:
    const char* p_;
    const char* end_;
    [[noreturn]] static void fail(const char* what) { throw std::runtime_error(std::string("JSON: ") + what); }
    void ws() {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    }
    bool lit(const char* s) {
        const size_t n = std::strlen(s);
        if ((size_t)(end_ - p_) >= n && std::memcmp(p_, s, n) == 0) { p_ += n; return true; }
        return false;
    }
    static void utf8(std::string& out, uint32_t c) {
        if (c < 0x80) out += (char)c;
        else if (c < 0x800) { out += (char)(0xC0 | (c >> 6)); out += (char)(0x80 | (c & 0x3F)); }
        else if (c < 0x10000) {
            out += (char)(0xE0 | (c >> 12)); out += (char)(0x80 | ((c >> 6) & 0x3F)); out += (char)(0x80 | (c & 0x3F));
        } else {
            out += (char)(0xF0 | (c >> 18)); out += (char)(0x80 | ((c >> 12) & 0x3F));
            out += (char)(0x80 | ((c >> 6) & 0x3F)); out += (char)(0x80 | (c & 0x3F));
        }
    }
    uint32_t hex4() {
        if (end_ - p_ < 4) fail("short \\u escape");
        uint32_t c = 0;
        for (int k = 0; k < 4; ++k) {
            const char h = *p_++;
            c <<= 4;
            if (h >= '0' && h <= '9') c |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f') c |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') c |= (uint32_t)(h - 'A' + 10);
            else fail("bad \\u escape");
        }
        return c;
    }
    std::string str() {
        ++p_;   // opening quote
        std::string out;
        for (;;) {
            if (p_ >= end_) fail("unterminated string");
            const char c = *p_++;
            if (c == '"') return out;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') { out += c; continue; }
            if (p_ >= end_) fail("unterminated escape");
            const char e = *p_++;
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && end_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
                        p_ += 2;
                        const uint32_t lo = hex4();
                        if (lo >= 0xDC00 && lo < 0xE000) cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                        else { utf8(out, cp); cp = lo; }
                    }
                    utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
    }
    double number() {
        const char* s = p_;
        if (p_ < end_ && *p_ == '-') ++p_;
        while (p_ < end_ && ((*p_ >= '0' && *p_ <= '9') || *p_ == '.' || *p_ == 'e' || *p_ == 'E' || *p_ == '+' ||
                             *p_ == '-'))
            ++p_;
        const std::string tok(s, p_);
        char* e = nullptr;
        const double v = std::strtod(tok.c_str(), &e);
        if (tok.empty() || e != tok.c_str() + tok.size()) fail("bad number");
        return v;
    }
    Json value(int depth) {
        if (depth > 256) fail("nesting too deep");
        ws();
        if (p_ >= end_) fail("unexpected end");
        Json v;
        const char c = *p_;
        if (c == '{') {
            v.kind = Json::Object;
            ++p_;
            ws();
            if (p_ < end_ && *p_ == '}') { ++p_; return v; }
            for (;;) {
                ws();
                if (p_ >= end_ || *p_ != '"') fail("object key expected");
                std::string k = str();
                ws();
                if (p_ >= end_ || *p_ != ':') fail("':' expected");
                ++p_;
                v.obj.emplace_back(std::move(k), value(depth + 1));
                ws();
                if (p_ < end_ && *p_ == ',') { ++p_; continue; }
                if (p_ < end_ && *p_ == '}') { ++p_; return v; }
                fail("',' or '}' expected");
            }
        }
        if (c == '[') {
            v.kind = Json::Array;
            ++p_;
            ws();
            if (p_ < end_ && *p_ == ']') { ++p_; return v; }
            for (;;) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (p_ < end_ && *p_ == ',') { ++p_; continue; }
                if (p_ < end_ && *p_ == ']') { ++p_; return v; }
                fail("',' or ']' expected");
            }
        }
        if (c == '"') { v.kind = Json::String; v.str = str(); return v; }
        if (lit("true")) { v.kind = Json::Bool; v.b = true; return v; }
        if (lit("false")) { v.kind = Json::Bool; v.b = false; return v; }
        if (lit("null")) return v;
        v.kind = Json::Number;
        v.num = number();
        return v;
    }
};

// ---------------------------------------------------------------- helpers
inline std::vector<uint8_t> base64_decode(const std::string& s) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0;
    for (char c : s) {
        if (c == '=') break;
        const int v = val(c);
        if (v < 0) {
            if (c == '\n' || c == '\r' || c == ' ') continue;
            throw std::runtime_error("bad base64 data");
        }
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)(acc >> bits));
        }
    }
    return out;
}

inline std::vector<uint8_t> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

inline std::string dir_of(const std::string& path) {
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string(".") : path.substr(0, k);
}

inline uint32_t le32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// glm::toMat4(glm::quat(w, x, y, z)) in float (column-major m[col][row]).
inline mat4 quat_to_mat4(float x, float y, float z, float w) {
    const float qxx = x * x, qyy = y * y, qzz = z * z;
    const float qxz = x * z, qxy = x * y, qyz = y * z;
    const float qwx = w * x, qwy = w * y, qwz = w * z;
    mat4 m;
    m[0][0] = 1.0f - 2.0f * (qyy + qzz);
    m[0][1] = 2.0f * (qxy + qwz);
    m[0][2] = 2.0f * (qxz - qwy);
    m[1][0] = 2.0f * (qxy - qwz);
    m[1][1] = 1.0f - 2.0f * (qxx + qzz);
    m[1][2] = 2.0f * (qyz + qwx);
    m[2][0] = 2.0f * (qxz + qwy);
    m[2][1] = 2.0f * (qyz - qwx);
    m[2][2] = 1.0f - 2.0f * (qxx + qyy);
    return m;
}

// glm mat4 * mat4 in float: column c = a[0] b[c][0] + a[1] b[c][1] + a[2] b[c][2]
// + a[3] b[c][3], summed left to right.
inline mat4 mat4_mul(const mat4& a, const mat4& b) {
    mat4 r;
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) {
            float acc = a[0][row] * b[c][0];
            acc = acc + a[1][row] * b[c][1];
            acc = acc + a[2][row] * b[c][2];
            acc = acc + a[3][row] * b[c][3];
            r[c][row] = acc;
        }
    return r;
}

// Inverse in double (Gauss-Jordan with partial pivoting), rounded to float.
inline mat4 inverse_f64(const mat4& m) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = (double)m[c][r];
            a[r][4 + c] = r == c ? 1.0 : 0.0;
        }
    for (int k = 0; k < 4; ++k) {
        int piv = k;
        for (int r = k + 1; r < 4; ++r)
            if (std::fabs(a[r][k]) > std::fabs(a[piv][k])) piv = r;
        if (a[piv][k] == 0.0) throw std::runtime_error("camera matrix is singular");
        if (piv != k)
            for (int c = 0; c < 8; ++c) std::swap(a[k][c], a[piv][c]);
        const double d = a[k][k];
        for (int c = 0; c < 8; ++c) a[k][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == k) continue;
            const double f = a[r][k];
            if (f == 0.0) continue;
            for (int c = 0; c < 8; ++c) a[r][c] -= f * a[k][c];
        }
    }
    mat4 out;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[c][r] = (float)a[r][4 + c];
    return out;
}

inline bool vec_of(const Json* j, size_t n, double* out) {
    if (!j || j->kind != Json::Array || j->arr.size() < n) return false;
    for (size_t k = 0; k < n; ++k) {
        if (j->arr[k].kind != Json::Number) throw std::runtime_error("non-numeric vector component");
        out[k] = j->arr[k].num;
    }
    return true;
}

}  // namespace importer_detail

class ModelImporter {
public:
    ModelImporter() = default;

    // Appends the file's textures, materials, meshes and cameras to *scene.
    bool loadGLTFfromFile(const char* file_path, BasicScene* scene) {
        error_.clear();
        if (!file_path || !scene) { error_ = "null argument"; return false; }
        try {
            BasicScene add;
            load(file_path, &add);
            for (auto& t : add.texture_entities) scene->addTexture(t);
            for (auto& m : add.material_entities) scene->addMaterial(m);
            for (auto& m : add.mesh_entities) scene->addMesh(m);
            for (auto& c : add.camera_entities) scene->addCamera(c);
            return true;
        } catch (const std::exception& e) {
            error_ = e.what();
            return false;
        }
    }
    const std::string& lastError() const { return error_; }

private:
    using Json = importer_detail::Json;
    std::string error_;
    Json doc_;
    std::vector<std::vector<uint8_t>> buffers_;
    std::string base_;

    [[noreturn]] static void fail(const std::string& what) { throw std::runtime_error(what); }
    // An index-valued member (buffer, bufferView, accessor, node, ...).
    static long long index_of(const Json& j) {
        if (j.kind != Json::Number || !(std::fabs(j.num) < 9.0e18)) fail("index is not a number");
        return (long long)j.num;
    }

    const Json& member(const Json& j, const char* key) const {
        const Json* v = j.get(key);
        if (!v) fail(std::string("missing '") + key + "'");
        return *v;
    }
    const Json& element(const char* array, long long i) const {
        const Json* a = doc_.get(array);
        const Json* v = (a && i >= 0) ? a->at((size_t)i) : nullptr;
        if (!v) fail(std::string(array) + " index out of range");
        return *v;
    }

    void load(const std::string& path, BasicScene* out) {
        using namespace importer_detail;
        std::vector<uint8_t> data = read_file(path);
        base_ = dir_of(path);
        std::vector<uint8_t> bin;
        bool have_bin = false;
        if (data.size() >= 12 && std::memcmp(data.data(), "glTF", 4) == 0) {
            if (le32(data.data() + 4) != 2) fail("glTF binary version not read");
            const size_t length = std::min<size_t>(le32(data.data() + 8), data.size());
            size_t pos = 12;
            bool have_json = false;
            while (pos + 8 <= length) {
                const size_t clen = le32(data.data() + pos), ctype = le32(data.data() + pos + 4);
                if (clen > data.size() - pos - 8) fail("glTF binary chunk runs past the file");
                const uint8_t* body = data.data() + pos + 8;
                if (ctype == 0x4E4F534Au) {
                    doc_ = JsonParser((const char*)body, clen).parse();
                    have_json = true;
                } else if (ctype == 0x004E4942u) {
                    bin.assign(body, body + clen);
                    have_bin = true;
                }
                pos += 8 + clen + ((4 - clen % 4) % 4);
            }
            if (!have_json) fail("glTF binary without a JSON chunk");
        } else {
            doc_ = JsonParser((const char*)data.data(), data.size()).parse();
        }
        buffers_.clear();
        if (const Json* bs = doc_.get("buffers")) {
            for (size_t i = 0; i < bs->size(); ++i) {
                const Json& b = bs->arr[i];
                const Json* uri = b.get("uri");
                if (!uri || uri->kind != Json::String) {
                    if (!have_bin) fail("buffer " + std::to_string(i) + " has no data");
                    buffers_.push_back(bin);
                } else if (uri->str.compare(0, 5, "data:") == 0) {
                    const size_t comma = uri->str.find(',');
                    if (comma == std::string::npos) fail("bad data URI");
                    buffers_.push_back(base64_decode(uri->str.substr(comma + 1)));
                } else {
                    buffers_.push_back(read_file(base_ + "/" + uri->str));
                }
            }
        }
        load_images(out);
        load_materials(out);
        load_nodes(out);
    }

    // [ptr, len) of a buffer view, and its byte stride (0: packed).
    const uint8_t* view_bytes(long long vi, size_t* len, size_t* stride) const {
        const Json& v = element("bufferViews", vi);
        const long long bi = index_of(member(v, "buffer"));
        if (bi < 0 || (size_t)bi >= buffers_.size()) fail("buffer index out of range");
        const std::vector<uint8_t>& buf = buffers_[(size_t)bi];
        const double off = v.number("byteOffset", 0.0), n = member(v, "byteLength").num;
        if (!(off >= 0) || !(n >= 0) || off + n > (double)buf.size()) fail("buffer view runs past its buffer");
        const long long st = v.integer("byteStride", 0);
        if (st != 0 && (st < 4 || st > 252)) fail("buffer view byteStride outside 4..252");
        *len = (size_t)n;
        *stride = (size_t)st;
        return buf.data() + (size_t)off;
    }

    // An accessor's elements as `want` floats each (components beyond the
    // accessor's count read 0), values cast from their component type.
    std::vector<float> accessor_f(long long ai, size_t want, size_t* count_out) const {
        const Json& a = element("accessors", ai);
        const int ct = (int)index_of(member(a, "componentType"));
        const std::string type = member(a, "type").str;
        const size_t nc = type == "SCALAR" ? 1 : type == "VEC2" ? 2 : type == "VEC3" ? 3 : type == "VEC4" ? 4
                        : type == "MAT4" ? 16 : 0;
        size_t csize = 0;
        switch (ct) {
            case 5120: case 5121: csize = 1; break;
            case 5122: case 5123: csize = 2; break;
            case 5125: case 5126: csize = 4; break;
            default: fail("unknown accessor component type");
        }
        if (nc == 0) fail("unknown accessor type " + type);
        const double cnt = member(a, "count").num;
        if (cnt < 0 || cnt > 1e9) fail("bad accessor count");
        const size_t count = (size_t)cnt;
        *count_out = count;
        std::vector<float> out(count * want, 0.0f);
        const Json* bv = a.get("bufferView");
        if (!bv) return out;   // no buffer view: zeros (glTF sparse accessors are not read)
        size_t len = 0, stride = 0;
        const uint8_t* base = view_bytes(index_of(*bv), &len, &stride);
        const double aoff = a.number("byteOffset", 0.0);
        if (!(aoff >= 0) || aoff > (double)len) fail("accessor offset outside its buffer view");
        const size_t off = (size_t)aoff, elem = csize * nc;
        if (stride == 0) stride = elem;
        if (count > 0 && off + (count - 1) * stride + elem > len) fail("accessor runs past its buffer view");
        for (size_t i = 0; i < count; ++i) {
            const uint8_t* e = base + off + i * stride;
            for (size_t k = 0; k < nc && k < want; ++k) {
                const uint8_t* c = e + k * csize;
                float v = 0.0f;
                switch (ct) {
                    case 5120: v = (float)(int8_t)c[0]; break;
                    case 5121: v = (float)c[0]; break;
                    case 5122: { int16_t x; std::memcpy(&x, c, 2); v = (float)x; break; }
                    case 5123: { uint16_t x; std::memcpy(&x, c, 2); v = (float)x; break; }
                    case 5125: { uint32_t x; std::memcpy(&x, c, 4); v = (float)x; break; }
                    default: std::memcpy(&v, c, 4); break;
                }
                out[i * want + k] = v;
            }
        }
        return out;
    }

    std::vector<int64_t> indices(long long ai) const {
        const Json& a = element("accessors", ai);
        const int ct = (int)index_of(member(a, "componentType"));
        const double cnt = member(a, "count").num;
        if (cnt < 0 || cnt > 1e9) fail("bad accessor count");
        const size_t count = (size_t)cnt;
        std::vector<int64_t> out(count, 0);
        const Json* bv = a.get("bufferView");
        if (!bv) return out;
        size_t csize = ct == 5121 ? 1 : ct == 5123 ? 2 : ct == 5125 ? 4 : 0;
        if (csize == 0) fail("index accessor is not an unsigned integer type");
        if (member(a, "type").str != "SCALAR") fail("index accessor is not SCALAR");
        size_t len = 0, stride = 0;
        const uint8_t* base = view_bytes(index_of(*bv), &len, &stride);
        const double aoff = a.number("byteOffset", 0.0);
        if (!(aoff >= 0) || aoff > (double)len) fail("accessor offset outside its buffer view");
        const size_t off = (size_t)aoff;
        if (stride == 0) stride = csize;
        if (count > 0 && off + (count - 1) * stride + csize > len) fail("accessor runs past its buffer view");
        for (size_t i = 0; i < count; ++i) {
            const uint8_t* c = base + off + i * stride;
            if (csize == 1) out[i] = c[0];
            else if (csize == 2) { uint16_t x; std::memcpy(&x, c, 2); out[i] = x; }
            else { uint32_t x; std::memcpy(&x, c, 4); out[i] = x; }
        }
        return out;
    }

    void load_images(BasicScene* out) const {
        using namespace importer_detail;
        const Json* ims = doc_.get("images");
        for (size_t i = 0; ims && i < ims->size(); ++i) {
            const Json& im = ims->arr[i];
            std::vector<uint8_t> bytes;
            if (const Json* bv = im.get("bufferView")) {
                size_t len = 0, stride = 0;
                const uint8_t* p = view_bytes(index_of(*bv), &len, &stride);
                bytes.assign(p, p + len);
            } else {
                const std::string uri = im.string("uri");
                if (uri.compare(0, 5, "data:") == 0) {
                    const size_t comma = uri.find(',');
                    if (comma == std::string::npos) fail("bad data URI");
                    bytes = base64_decode(uri.substr(comma + 1));
                } else {
                    bytes = read_file(base_ + "/" + uri);
                }
            }
            static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
            if (bytes.size() < 8 || std::memcmp(bytes.data(), sig, 8) != 0)
                fail("image " + std::to_string(i) + ": only PNG images are decoded here");
            int32_t w = 0, h = 0;
            if (kpt_png_decode_rgb8(bytes.data(), (int64_t)bytes.size(), &w, &h, nullptr, 0) != KPT_OK)
                fail("image " + std::to_string(i) + ": PNG not decoded");
            std::vector<uint8_t> rgb((size_t)w * h * 3);
            if (kpt_png_decode_rgb8(bytes.data(), (int64_t)bytes.size(), &w, &h, rgb.data(), (int64_t)rgb.size()) !=
                KPT_OK)
                fail("image " + std::to_string(i) + ": PNG not decoded");
            out->addTexture(TextureSceneEntity(im.string("name"), rgb.data(), (uint32_t)w, (uint32_t)h, 3));
        }
    }

    int tex_source(const Json* ref) const {
        if (!ref || ref->integer("index", -1) < 0) return -1;
        return (int)element("textures", ref->integer("index", -1)).integer("source", -1);
    }

    void load_materials(BasicScene* out) const {
        using namespace importer_detail;
        const Json* ms = doc_.get("materials");
        const Json empty;
        for (size_t i = 0; ms && i < ms->size(); ++i) {
            const Json& m = ms->arr[i];
            const Json* pbr = m.get("pbrMetallicRoughness");
            if (!pbr) pbr = &empty;
            double bc[4] = {1.0, 1.0, 1.0, 1.0};
            vec_of(pbr->get("baseColorFactor"), 3, bc);
            const Json* ext = m.get("extensions");
            if (!ext) ext = &empty;
            const Json* tr = ext->get("KHR_materials_transmission");
            if (!tr) tr = &empty;
            const Json* ior = ext->get("KHR_materials_ior");
            const Json* es = ext->get("KHR_materials_emissive_strength");
            double em[3] = {0.0, 0.0, 0.0};
            vec_of(m.get("emissiveFactor"), 3, em);
            const Json* nt = m.get("normalTexture");
            out->addMaterial(MaterialSceneEntity(
                m.string("name"), tex_source(pbr->get("baseColorTexture")),
                vec3((float)bc[0], (float)bc[1], (float)bc[2]), tex_source(pbr->get("metallicRoughnessTexture")),
                (float)pbr->number("metallicFactor", 1.0), (float)pbr->number("roughnessFactor", 1.0),
                tex_source(tr->get("transmissionTexture")), (float)tr->number("transmissionFactor", 0.0),
                (float)(ior ? ior->number("ior", 1.45) : 1.45), tex_source(m.get("emissiveTexture")),
                vec3((float)em[0], (float)em[1], (float)em[2]),
                (float)(es ? es->number("emissiveStrength", 0.0) : 0.0), tex_source(nt),
                (float)(nt ? nt->number("scale", 1.0) : 1.0)));
        }
        if (!ms || ms->size() == 0) {
            // the Python importer's default material (scenes.material())
            out->addMaterial(MaterialSceneEntity("default", -1, vec3(0.8f), -1, 0.0f, 0.5f, -1, 0.0f, 1.45f, -1,
                                                 vec3(1.0f), 0.0f, -1, 1.0f));
        }
    }

    static mat4 translation(const importer_detail::Json& node) {
        mat4 t;
        double tr[3];
        if (importer_detail::vec_of(node.get("translation"), 3, tr)) {
            t[3][0] = (float)tr[0]; t[3][1] = (float)tr[1]; t[3][2] = (float)tr[2];
        }
        return t;
    }
    static bool matrix_of(const importer_detail::Json& node, mat4* m) {
        double v[16];
        if (!importer_detail::vec_of(node.get("matrix"), 16, v)) return false;
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) (*m)[c][r] = (float)v[4 * c + r];
        return true;
    }

    // parseMesh's model matrix (:155-198).
    static mat4 node_model(const importer_detail::Json& node) {
        using namespace importer_detail;
        mat4 m;
        if (matrix_of(node, &m)) return m;
        mat4 s, r;
        double v[4];
        if (vec_of(node.get("scale"), 3, v)) { s[0][0] = (float)v[0]; s[1][1] = (float)v[1]; s[2][2] = (float)v[2]; }
        if (vec_of(node.get("rotation"), 4, v)) r = quat_to_mat4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
        return mat4_mul(mat4_mul(translation(node), r), s);
    }

    // parseCamera's view matrix (:380-410), inverted in double.
    static mat4 camera_view(const importer_detail::Json& node) {
        using namespace importer_detail;
        mat4 model;
        if (!matrix_of(node, &model)) {
            mat4 r;
            double v[4];
            if (vec_of(node.get("rotation"), 4, v)) {
                float x = (float)v[0], y = (float)v[1], z = (float)v[2], w = (float)v[3];
                const float ss = ((x * x + y * y) + z * z) + w * w;
                const float ln = (float)std::sqrt((double)ss);   // glm::normalize(quat)
                if (ln > 0.0f) {
                    const float inv = 1.0f / ln;
                    x = x * inv; y = y * inv; z = z * inv; w = w * inv;
                }
                r = quat_to_mat4(x, y, z, w);
            }
            model = mat4_mul(translation(node), r);
        }
        return inverse_f64(model);
    }

    void load_nodes(BasicScene* out) const {
        const Json* scenes = doc_.get("scenes");
        const Json* s0 = scenes ? scenes->at(0) : nullptr;
        const Json* nodes = s0 ? s0->get("nodes") : nullptr;   // scene 0, as the reference
        size_t meshes = 0;
        for (size_t k = 0; nodes && k < nodes->size(); ++k) {
            const Json& node = element("nodes", index_of(nodes->arr[k]));
            if (node.integer("camera", -1) >= 0) {
                const Json& c = element("cameras", node.integer("camera", -1));
                const Json* persp = c.get("perspective");
                const double yfov = persp ? persp->number("yfov", 45.0 * 3.14159265358979323846 / 180.0)
                                          : 45.0 * 3.14159265358979323846 / 180.0;
                out->addCamera(CameraSceneEntity(c.string("name"), camera_view(node), (float)yfov));
            }
            if (node.integer("mesh", -1) < 0) continue;
            const Json& mesh = element("meshes", node.integer("mesh", -1));
            MeshSceneEntity me(mesh.string("name"), node_model(node));
            const Json& prims = member(mesh, "primitives");
            for (size_t p = 0; p < prims.size(); ++p) {
                const Json& prim = prims.arr[p];
                if (prim.integer("mode", 4) != 4) fail("only triangle primitives are read");
                const Json& at = member(prim, "attributes");
                size_t np = 0, nn = 0, nu = 0;
                const std::vector<float> P = accessor_f(index_of(member(at, "POSITION")), 3, &np);
                const std::vector<float> N = accessor_f(index_of(member(at, "NORMAL")), 3, &nn);
                std::vector<float> U;
                if (const Json* uv = at.get("TEXCOORD_0")) U = accessor_f(index_of(*uv), 2, &nu);
                else U.assign(np * 2, 0.0f);
                if (nn < np || (at.get("TEXCOORD_0") && nu < np)) fail("vertex attribute counts differ");
                std::vector<int64_t> idx;
                if (const Json* ia = prim.get("indices")) idx = indices(index_of(*ia));
                else for (size_t i = 0; i < np; ++i) idx.push_back((int64_t)i);
                const int32_t mat = (int32_t)prim.integer("material", 0);
                for (size_t t = 0; t + 2 < idx.size(); t += 3) {
                    vec3 pv[3], nv[3];
                    vec2 tv[3];
                    for (int c = 0; c < 3; ++c) {
                        const int64_t i = idx[t + c];
                        if (i < 0 || (size_t)i >= np) fail("vertex index out of range");
                        pv[c] = vec3(P[3 * i], P[3 * i + 1], P[3 * i + 2]);
                        nv[c] = vec3(N[3 * i], N[3 * i + 1], N[3 * i + 2]);
                        tv[c].x = U[2 * i];
                        tv[c].y = U[2 * i + 1];
                    }
                    me.addShape(TriangleSceneEntity(pv[0], pv[1], pv[2], nv[0], nv[1], nv[2], tv[0], tv[1], tv[2], mat));
                }
            }
            out->addMesh(me);
            ++meshes;
        }
        if (meshes == 0) fail("the first scene holds no mesh node");
    }
};

}  // namespace KittlesPT
