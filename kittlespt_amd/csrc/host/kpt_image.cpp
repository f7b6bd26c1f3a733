// kpt_image.cpp — image-data helpers of the C-ABI (host code).
//
// PNG scanline unfiltering (the PNG specification, section 9: filter types
// None, Sub, Up, Average, Paeth on the bytes of one row, each predicting from
// the byte bpp to the left, the byte above and the byte above-left).  The
// glTF importer decodes texture images as the reference's stb_image does
// (SampleApp's model importer, model_importer.hpp); the Sub, Average and
// Paeth filters run along a row byte by byte, which is what this does in C++.
#include <cstdint>
#include <cstdlib>

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
