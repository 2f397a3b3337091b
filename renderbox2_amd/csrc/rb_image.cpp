// rb_image.cpp -- image writers for rendered tiles (ImageWriter::write,
// src/io/imagewriter.cpp:34-105).
//
// The reference picks the format by suffix and hands the pixels to vendored
// encoders (miniexr for half-float EXR, jo_jpeg for 8-bit JPEG); its EXR path
// writes the converted half pixels instead of the encoded file.  Written
// here from the file format specifications:
//  * .exr : OpenEXR scanline file, uncompressed, HALF channels B, G, R;
//  * .pfm : portable float map, float32 RGB, bottom row first;
//  * .ppm : binary PPM, 8-bit RGB;
//  * .png : 8-bit RGB, zlib stream of stored (uncompressed) deflate blocks.
// 8-bit values use the reference's conversion (ImageWriter::to_byte,
// src/io/imagewriter.h:48): clamp to [0, 1], gamma 1/2.2, round.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "rb_internal.h"

namespace rb {
namespace {

// binary32 -> binary16, round to nearest even; overflow to infinity, NaN kept.
uint16_t to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t e = (x >> 23) & 0xFFu;
    uint32_t m = x & 0x7FFFFFu;
    if (e == 0xFFu) return (uint16_t)(sign | 0x7C00u | (m ? 0x200u | (m >> 13) : 0u));
    const int E = (int)e - 127 + 15;
    if (E >= 31) return (uint16_t)(sign | 0x7C00u);
    if (E <= 0) {  // subnormal half (or zero)
        if (E < -10) return (uint16_t)sign;
        m |= 0x800000u;
        const int shift = 14 - E;  // 14..24
        uint32_t h = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)E << 10) | (m >> 13);
    const uint32_t rem = m & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;  // may carry into the exponent: correct
    return (uint16_t)(sign | h);
}

unsigned char to_byte(float v) {
    const float c = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    return (unsigned char)(int)(std::pow(c, 1.0f / 2.2f) * 255 + .5);  // float product, double rounding add
}

uint32_t crc32(const unsigned char* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t r = i;
            for (int k = 0; k < 8; k++) r = (r & 1u) ? 0xEDB88320u ^ (r >> 1) : r >> 1;
            table[i] = r;
        }
        init = true;
    }
    for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c;
}

void put32be(std::string& s, uint32_t v) {
    s += (char)(v >> 24);
    s += (char)(v >> 16);
    s += (char)(v >> 8);
    s += (char)v;
}

template <class T>
void put_le(std::string& s, T v) {
    s.append(reinterpret_cast<const char*>(&v), sizeof(T));
}

void png_chunk(std::string& out, const char* type, const std::string& data) {
    put32be(out, (uint32_t)data.size());
    std::string td = std::string(type, 4) + data;
    out += td;
    put32be(out, crc32(reinterpret_cast<const unsigned char*>(td.data()), td.size()) ^ 0xFFFFFFFFu);
}

std::string encode_png(const float* rgb, uint32_t w, uint32_t h) {
    std::string raw;
    raw.reserve((size_t)h * (3 * (size_t)w + 1));
    for (uint32_t y = 0; y < h; y++) {
        raw += '\0';  // filter: none
        for (uint32_t x = 0; x < 3 * w; x++) raw += (char)to_byte(rgb[(size_t)y * 3 * w + x]);
    }
    std::string z;
    z += (char)0x78;
    z += (char)0x01;
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + len == raw.size();
        z += (char)(last ? 1 : 0);
        z += (char)(len & 0xFF);
        z += (char)(len >> 8);
        z += (char)(~len & 0xFF);
        z += (char)((~len >> 8) & 0xFF);
        z.append(raw, pos, len);
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (unsigned char c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put32be(z, (b << 16) | a);
    std::string ihdr;
    put32be(ihdr, w);
    put32be(ihdr, h);
    ihdr += (char)8;  // bit depth
    ihdr += (char)2;  // colour type: RGB
    ihdr += std::string(3, '\0');  // deflate, adaptive filtering, no interlace
    std::string out("\x89PNG\r\n\x1a\n", 8);
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", std::string());
    return out;
}

void exr_attr(std::string& s, const char* name, const char* type, const std::string& value) {
    s += name;
    s += '\0';
    s += type;
    s += '\0';
    put_le<int32_t>(s, (int32_t)value.size());
    s += value;
}

std::string encode_exr(const float* rgb, uint32_t w, uint32_t h) {
    std::string s;
    put_le<uint32_t>(s, 20000630u);  // magic 0x762f3101
    put_le<uint32_t>(s, 2u);         // version 2, scanline
    std::string ch;
    for (const char* c : {"B", "G", "R"}) {  // channel list sorted by name
        ch += c;
        ch += '\0';
        put_le<int32_t>(ch, 1);  // HALF
        ch += std::string(4, '\0');  // pLinear + reserved
        put_le<int32_t>(ch, 1);
        put_le<int32_t>(ch, 1);
    }
    ch += '\0';
    exr_attr(s, "channels", "chlist", ch);
    exr_attr(s, "compression", "compression", std::string(1, '\0'));
    std::string box;
    put_le<int32_t>(box, 0);
    put_le<int32_t>(box, 0);
    put_le<int32_t>(box, (int32_t)w - 1);
    put_le<int32_t>(box, (int32_t)h - 1);
    exr_attr(s, "dataWindow", "box2i", box);
    exr_attr(s, "displayWindow", "box2i", box);
    exr_attr(s, "lineOrder", "lineOrder", std::string(1, '\0'));
    std::string f1;
    put_le<float>(f1, 1.0f);
    exr_attr(s, "pixelAspectRatio", "float", f1);
    std::string v2;
    put_le<float>(v2, 0.0f);
    put_le<float>(v2, 0.0f);
    exr_attr(s, "screenWindowCenter", "v2f", v2);
    exr_attr(s, "screenWindowWidth", "float", f1);
    s += '\0';
    const size_t line_bytes = (size_t)w * 3 * 2;
    uint64_t off = s.size() + 8ull * h;
    for (uint32_t y = 0; y < h; y++) {
        put_le<uint64_t>(s, off);
        off += 8 + line_bytes;
    }
    for (uint32_t y = 0; y < h; y++) {
        put_le<int32_t>(s, (int32_t)y);
        put_le<int32_t>(s, (int32_t)line_bytes);
        for (int c = 2; c >= 0; c--)  // B, G, R
            for (uint32_t x = 0; x < w; x++) put_le<uint16_t>(s, to_half(rgb[((size_t)y * w + x) * 3 + c]));
    }
    return s;
}

bool ends_with(const std::string& s, const char* suf) {
    const size_t n = std::strlen(suf);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; i++)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
    return true;
}

}  // namespace
}  // namespace rb

using namespace rb;

extern "C" {

rb_status rb_image_write(const char* path, const float* rgb, uint32_t w, uint32_t h) {
    if (!path || !rgb) return fail(RB_ERR_INVALID, "NULL argument");
    if (w == 0 || h == 0 || (uint64_t)w * h > (1ull << 30)) return fail(RB_ERR_INVALID, "bad image size %ux%u", w, h);
    const std::string p(path);
    std::string data;
    try {
        if (ends_with(p, ".exr")) {
            data = encode_exr(rgb, w, h);
        } else if (ends_with(p, ".png")) {
            data = encode_png(rgb, w, h);
        } else if (ends_with(p, ".ppm")) {
            data = "P6\n" + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
            data.reserve(data.size() + (size_t)w * h * 3);
            for (size_t i = 0; i < (size_t)w * h * 3; i++) data += (char)to_byte(rgb[i]);
        } else if (ends_with(p, ".pfm")) {
            data = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.0\n";  // little-endian
            for (uint32_t y = h; y-- > 0;)
                data.append(reinterpret_cast<const char*>(rgb + (size_t)y * w * 3), (size_t)w * 12);
        } else {
            return fail(RB_ERR_UNSUPPORTED, "%s: unknown image suffix (exr, png, ppm, pfm)", path);
        }
    } catch (...) {
        return fail(RB_ERR_OOM, "image encoding failed (host allocation)");
    }
    std::ofstream out(p, std::ios::binary | std::ios::trunc);
    if (!out) return fail(RB_ERR_INVALID, "cannot open %s", path);
    out.write(data.data(), (std::streamsize)data.size());
    return out ? RB_OK : fail(RB_ERR_INVALID, "write to %s failed", path);
}

}  // extern "C"
