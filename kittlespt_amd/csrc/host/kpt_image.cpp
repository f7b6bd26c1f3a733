// kpt_image.cpp — image-data helpers of the C-ABI (host code): PNG decoding
// for the glTF importers (Python and C++), and PNG/PFM/OpenEXR files of
// read-back frames.
//
// PNG scanline unfiltering (the PNG specification, section 9: filter types
// None, Sub, Up, Average, Paeth on the bytes of one row, each predicting from
// the byte bpp to the left, the byte above and the byte above-left).  The
// glTF importer decodes texture images as the reference's stb_image does
// (SampleApp's model importer, model_importer.hpp); the Sub, Average and
// Paeth filters run along a row byte by byte, which is what this does in C++.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <zlib.h>

#include "../../../include/kpt_abi.h"

namespace {

inline uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    return (uint8_t)(pb <= pc ? b : c);
}

}  // namespace

extern "C" int kpt_png_unfilter(const uint8_t* raw, int64_t height, int64_t stride, int32_t bpp, uint8_t* out) {
    if (height < 0 || stride < 0 || bpp < 1 || bpp > 8 || (height > 0 && stride > 0 && (!raw || !out)))
        return KPT_ERR_INVALID_ARG;
    for (int64_t y = 0; y < height; ++y) {
        const uint8_t* line = raw + y * (stride + 1);
        const uint8_t f = line[0];
        ++line;
        uint8_t* cur = out + y * stride;
        const uint8_t* prev = y > 0 ? out + (y - 1) * stride : nullptr;
        for (int64_t x = 0; x < stride; ++x) {
            const int a = x >= bpp ? cur[x - bpp] : 0;
            const int b = prev ? prev[x] : 0;
            const int c = (prev && x >= bpp) ? prev[x - bpp] : 0;
            int pred;
            switch (f) {
                case 0: pred = 0; break;
                case 1: pred = a; break;
                case 2: pred = b; break;
                case 3: pred = (a + b) >> 1; break;
                case 4: pred = paeth(a, b, c); break;
                default: return KPT_ERR_INVALID_ARG;
            }
            cur[x] = (uint8_t)(line[x] + pred);
        }
    }
    return KPT_OK;
}

// ---------------------------------------------------------------- PNG decode
// The stb_image call of the reference's importer (stbi_load_from_memory with
// 3 requested channels, model_importer.hpp:287-296) for the PNG images glTF
// files carry: 8-bit, non-interlaced grey, RGB, palette, grey-alpha and RGBA
// images become RGB8 (grey replicated, alpha dropped, palette looked up).
// The IDAT stream is inflated with zlib and unfiltered by kpt_png_unfilter.
namespace {

inline uint32_t be32(const uint8_t* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}

struct PngInfo {
    uint32_t w = 0, h = 0;
    int ctype = -1, chans = 0;
    std::vector<uint8_t> idat, plte;
};

// Walks the chunks; false on a malformed stream or an image this decoder
// does not read (the caller reports KPT_ERR_UNSUPPORTED for the latter).
int png_parse(const uint8_t* d, int64_t n, PngInfo* info) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (n < 8 || std::memcmp(d, sig, 8) != 0) return KPT_ERR_INVALID_ARG;
    int64_t pos = 8;
    bool have_hdr = false;
    int depth = 0, interlace = 0;
    while (pos + 12 <= n) {
        const uint32_t len = be32(d + pos);
        if ((int64_t)len > n - pos - 12) return KPT_ERR_INVALID_ARG;
        const uint8_t* tag = d + pos + 4;
        const uint8_t* body = d + pos + 8;
        if (std::memcmp(tag, "IHDR", 4) == 0) {
            if (len < 13) return KPT_ERR_INVALID_ARG;
            info->w = be32(body);
            info->h = be32(body + 4);
            depth = body[8];
            info->ctype = body[9];
            interlace = body[12];
            have_hdr = true;
        } else if (std::memcmp(tag, "PLTE", 4) == 0) {
            info->plte.assign(body, body + len);
        } else if (std::memcmp(tag, "IDAT", 4) == 0) {
            info->idat.insert(info->idat.end(), body, body + len);
        } else if (std::memcmp(tag, "IEND", 4) == 0) {
            break;
        }
        pos += 12 + (int64_t)len;
    }
    if (!have_hdr) return KPT_ERR_INVALID_ARG;
    if (depth != 8 || interlace != 0) return KPT_ERR_UNSUPPORTED;
    switch (info->ctype) {
        case 0: info->chans = 1; break;
        case 2: info->chans = 3; break;
        case 3: info->chans = 1; break;
        case 4: info->chans = 2; break;
        case 6: info->chans = 4; break;
        default: return KPT_ERR_UNSUPPORTED;
    }
    // at most 2^28 pixels: the inflated stream (<= 4 bytes + 1 per pixel row
    // byte) then fits zlib's 32-bit buffer lengths
    if (info->w == 0 || info->h == 0 || (uint64_t)info->w * info->h > (1ull << 28)) return KPT_ERR_UNSUPPORTED;
    return KPT_OK;
}

}  // namespace

extern "C" int kpt_png_decode_rgb8(const uint8_t* data, int64_t n, int32_t* width, int32_t* height, uint8_t* rgb,
                                   int64_t capacity) {
    if (!data || n < 0 || !width || !height) return KPT_ERR_INVALID_ARG;
    PngInfo info;
    int rc = png_parse(data, n, &info);
    if (rc != KPT_OK) return rc;
    *width = (int32_t)info.w;
    *height = (int32_t)info.h;
    if (!rgb) return KPT_OK;
    const int64_t npx = (int64_t)info.w * info.h;
    if (capacity < npx * 3) return KPT_ERR_INVALID_ARG;
    if (info.ctype == 3 && info.plte.size() < 3) return KPT_ERR_INVALID_ARG;
    const int64_t stride = (int64_t)info.w * info.chans;
    const int64_t raw_len = (int64_t)info.h * (stride + 1);
    std::vector<uint8_t> raw((size_t)raw_len);
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return KPT_ERR_OUT_OF_MEMORY;
    if (info.idat.size() > 0xFFFFFFFFull) return KPT_ERR_UNSUPPORTED;
    zs.next_in = info.idat.data();
    zs.avail_in = (uInt)info.idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const uint64_t produced = zs.total_out;
    inflateEnd(&zs);
    if (zr != Z_STREAM_END || produced != (uint64_t)raw_len) return KPT_ERR_INVALID_ARG;
    std::vector<uint8_t> px((size_t)(stride * info.h));
    rc = kpt_png_unfilter(raw.data(), info.h, stride, info.chans, px.data());
    if (rc != KPT_OK) return rc;
    const size_t ncol = info.plte.size() / 3;
    for (int64_t i = 0; i < npx; ++i) {
        const uint8_t* s = px.data() + i * info.chans;
        uint8_t* o = rgb + i * 3;
        switch (info.ctype) {
            case 0: case 4: o[0] = o[1] = o[2] = s[0]; break;
            case 3: {
                if (s[0] >= ncol) return KPT_ERR_INVALID_ARG;
                const uint8_t* c = info.plte.data() + 3 * (size_t)s[0];
                o[0] = c[0]; o[1] = c[1]; o[2] = c[2];
                break;
            }
            default: o[0] = s[0]; o[1] = s[1]; o[2] = s[2]; break;
        }
    }
    return KPT_OK;
}

// ---------------------------------------------------------------- frame files
// The reference shows its frames through GL textures (renderer.cu:197-205);
// these write read-back frames to files.  Rows are taken as the frame
// buffers keep them (row 0 = bottom, the GL convention) and flipped to the
// top-down order of PNG and OpenEXR when flip_y is set; PFM is bottom-up.
namespace {

struct File {
    FILE* f;
    explicit File(const char* path) : f(std::fopen(path, "wb")) {}
    ~File() { if (f) std::fclose(f); }
    bool put(const void* p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }
};

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24)); v.push_back((uint8_t)(x >> 16)); v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}
template <typename T>
void put_le(std::vector<uint8_t>& v, T x) {
    uint8_t b[sizeof(T)];
    std::memcpy(b, &x, sizeof(T));
    v.insert(v.end(), b, b + sizeof(T));
}

bool png_chunk(File& f, const char* tag, const uint8_t* body, size_t n) {
    std::vector<uint8_t> c;
    put_be32(c, (uint32_t)n);
    c.insert(c.end(), tag, tag + 4);
    c.insert(c.end(), body, body + n);
    uLong crc = crc32(0L, Z_NULL, 0);
    crc = crc32(crc, c.data() + 4, (uInt)(n + 4));
    put_be32(c, (uint32_t)crc);
    return f.put(c.data(), c.size());
}

// floor(v * 255 + 0.5) clamped to [0, 255]; NaN -> 0, +inf -> 255, -inf -> 0.
uint8_t to_byte(float v) {
    if (v != v) v = 0.0f;
    if (v > 3.0e38f) v = 1.0f;
    if (v < -3.0e38f) v = 0.0f;
    float s = v * 255.0f;
    s = s + 0.5f;
    const float fl = std::floor(s);
    return (uint8_t)(fl < 0.0f ? 0.0f : (fl > 255.0f ? 255.0f : fl));
}

bool frame_args(const char* path, const float* px, int32_t w, int32_t h, int32_t channels, int min_ch, int max_ch) {
    return path && px && w > 0 && h > 0 && channels >= min_ch && channels <= max_ch &&
           (int64_t)w * h <= (1ll << 30);
}

}  // namespace

extern "C" int kpt_write_png_rgba8(const char* path, const float* px, int32_t width, int32_t height, int32_t channels,
                                   int32_t flip_y) {
    if (!frame_args(path, px, width, height, channels, 3, 4)) return KPT_ERR_INVALID_ARG;
    const int64_t row = 1 + 4 * (int64_t)width;
    std::vector<uint8_t> raw((size_t)(row * height));
    for (int32_t y = 0; y < height; ++y) {
        const int32_t sy = flip_y ? height - 1 - y : y;
        uint8_t* o = raw.data() + y * row;
        o[0] = 0;
        for (int32_t x = 0; x < width; ++x) {
            const float* s = px + ((int64_t)sy * width + x) * channels;
            for (int c = 0; c < 4; ++c) o[1 + 4 * x + c] = to_byte(c < channels ? s[c] : 1.0f);
        }
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return KPT_ERR_OUT_OF_MEMORY;
    File f(path);
    if (!f.f) return KPT_ERR_INVALID_ARG;
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)width);
    put_be32(ihdr, (uint32_t)height);
    const uint8_t rest[5] = {8, 6, 0, 0, 0};
    ihdr.insert(ihdr.end(), rest, rest + 5);
    if (!f.put(sig, 8) || !png_chunk(f, "IHDR", ihdr.data(), ihdr.size()) || !png_chunk(f, "IDAT", z.data(), zlen) ||
        !png_chunk(f, "IEND", nullptr, 0))
        return KPT_ERR_INVALID_ARG;
    return KPT_OK;
}

extern "C" int kpt_write_pfm(const char* path, const float* px, int32_t width, int32_t height, int32_t channels) {
    if (!frame_args(path, px, width, height, channels, 3, 4)) return KPT_ERR_INVALID_ARG;
    File f(path);
    if (!f.f) return KPT_ERR_INVALID_ARG;
    const std::string hdr = "PF\n" + std::to_string(width) + " " + std::to_string(height) + "\n-1.0\n";
    if (!f.put(hdr.data(), hdr.size())) return KPT_ERR_INVALID_ARG;
    std::vector<float> line((size_t)width * 3);
    for (int32_t y = 0; y < height; ++y) {
        for (int32_t x = 0; x < width; ++x)
            for (int c = 0; c < 3; ++c) line[(size_t)x * 3 + c] = px[((int64_t)y * width + x) * channels + c];
        if (!f.put(line.data(), line.size() * sizeof(float))) return KPT_ERR_INVALID_ARG;
    }
    return KPT_OK;
}

// Single-part scanline OpenEXR, no compression, FLOAT channels A, B, G, R
// (the format keeps channels in alphabetical order), increasing-Y lines.
extern "C" int kpt_write_exr(const char* path, const float* px, int32_t width, int32_t height, int32_t channels,
                             int32_t flip_y) {
    if (!frame_args(path, px, width, height, channels, 3, 4)) return KPT_ERR_INVALID_ARG;
    std::vector<uint8_t> h;
    put_le<int32_t>(h, 20000630);
    put_le<int32_t>(h, 2);
    auto attr = [&h](const char* name, const char* type, const std::vector<uint8_t>& data) {
        h.insert(h.end(), name, name + std::strlen(name) + 1);
        h.insert(h.end(), type, type + std::strlen(type) + 1);
        put_le<int32_t>(h, (int32_t)data.size());
        h.insert(h.end(), data.begin(), data.end());
    };
    std::vector<uint8_t> ch;
    for (const char* c : {"A", "B", "G", "R"}) {
        ch.push_back((uint8_t)c[0]);
        ch.push_back(0);
        put_le<int32_t>(ch, 2);                    // FLOAT
        ch.push_back(0); ch.push_back(0); ch.push_back(0); ch.push_back(0);   // pLinear, reserved
        put_le<int32_t>(ch, 1);                    // x sampling
        put_le<int32_t>(ch, 1);                    // y sampling
    }
    ch.push_back(0);
    std::vector<uint8_t> box;
    put_le<int32_t>(box, 0); put_le<int32_t>(box, 0); put_le<int32_t>(box, width - 1); put_le<int32_t>(box, height - 1);
    std::vector<uint8_t> one, centre, zero1(1, 0);
    put_le<float>(one, 1.0f);
    put_le<float>(centre, 0.0f); put_le<float>(centre, 0.0f);
    attr("channels", "chlist", ch);
    attr("compression", "compression", zero1);
    attr("dataWindow", "box2i", box);
    attr("displayWindow", "box2i", box);
    attr("lineOrder", "lineOrder", zero1);
    attr("pixelAspectRatio", "float", one);
    attr("screenWindowCenter", "v2f", centre);
    attr("screenWindowWidth", "float", one);
    h.push_back(0);
    const int64_t line_bytes = 4 * (int64_t)width * 4;
    const uint64_t first = h.size() + 8 * (uint64_t)height;
    for (int32_t y = 0; y < height; ++y) put_le<uint64_t>(h, first + (uint64_t)y * (8 + line_bytes));
    File f(path);
    if (!f.f) return KPT_ERR_INVALID_ARG;
    if (!f.put(h.data(), h.size())) return KPT_ERR_INVALID_ARG;
    std::vector<uint8_t> line;
    for (int32_t y = 0; y < height; ++y) {
        const int32_t sy = flip_y ? height - 1 - y : y;
        line.clear();
        put_le<int32_t>(line, y);
        put_le<int32_t>(line, (int32_t)line_bytes);
        for (int c : {3, 2, 1, 0})
            for (int32_t x = 0; x < width; ++x) {
                const float* s = px + ((int64_t)sy * width + x) * channels;
                put_le<float>(line, c < channels ? s[c] : 1.0f);
            }
        if (!f.put(line.data(), line.size())) return KPT_ERR_INVALID_ARG;
    }
    return KPT_OK;
}
