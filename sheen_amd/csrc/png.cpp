// png.cpp — sRGB encoding and an 8-bit RGB PNG writer (zlib deflate).
//
// The reference writes its PNG through stb_image_write after the sRGB curve
// of png_color.h:11-21; the byte stream here comes from zlib, so files differ
// in their compressed bytes while the decoded pixels are the same.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "host_util.h"
#include "sheen_c_api.h"

unsigned char host_srgb_byte(float c) {
    float e;
    if (!(c > 0.0f)) e = 0.0f;
    else {
        if (c > 1.0f) c = 1.0f;
        e = (c <= 0.0031308f) ? c * 12.92f : 1.055f * std::pow(c, 1.0f / 2.4f) - 0.055f;
    }
    float v = e * 255.0f + 0.5f;
    return (unsigned char)(v > 255.0f ? 255.0f : v);
}

extern "C" void sheen_encode_srgb(const float* rgba, int width, int height, unsigned char* out_rgb) {
    const size_t n = (size_t)width * height;
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) out_rgb[3 * i + c] = host_srgb_byte(rgba[4 * i + c]);
}

static void put_be32(std::vector<unsigned char>& b, uint32_t v) {
    b.push_back((unsigned char)(v >> 24)); b.push_back((unsigned char)(v >> 16));
    b.push_back((unsigned char)(v >> 8)); b.push_back((unsigned char)v);
}
static void put_chunk(std::vector<unsigned char>& out, const char* type, const unsigned char* data, size_t n) {
    put_be32(out, (uint32_t)n);
    size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    if (n) out.insert(out.end(), data, data + n);
    uLong crc = crc32(0L, Z_NULL, 0);
    crc = crc32(crc, out.data() + start, (uInt)(n + 4));
    put_be32(out, (uint32_t)crc);
}

extern "C" int sheen_write_png(const char* path, int width, int height, const unsigned char* rgb) {
    if (!path || !rgb || width <= 0 || height <= 0) return -1;
    const size_t row = (size_t)width * 3;
    std::vector<unsigned char> raw((row + 1) * height);
    for (int y = 0; y < height; y++) {
        raw[(row + 1) * y] = 0;  // filter: none
        std::memcpy(&raw[(row + 1) * y + 1], rgb + row * y, row);
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<unsigned char> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK) return -1;
    std::vector<unsigned char> out;
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    out.insert(out.end(), sig, sig + 8);
    unsigned char ihdr[13];
    uint32_t w = (uint32_t)width, h = (uint32_t)height;
    ihdr[0] = w >> 24; ihdr[1] = w >> 16; ihdr[2] = w >> 8; ihdr[3] = w;
    ihdr[4] = h >> 24; ihdr[5] = h >> 16; ihdr[6] = h >> 8; ihdr[7] = h;
    ihdr[8] = 8; ihdr[9] = 2; ihdr[10] = 0; ihdr[11] = 0; ihdr[12] = 0;
    put_chunk(out, "IHDR", ihdr, 13);
    put_chunk(out, "IDAT", z.data(), zn);
    put_chunk(out, "IEND", nullptr, 0);
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    size_t wr = std::fwrite(out.data(), 1, out.size(), f);
    int rc = std::fclose(f);
    return (wr == out.size() && rc == 0) ? 0 : -1;
}
