// rb_image.cpp -- image writers for rendered tiles (ImageWriter::write,
// src/io/imagewriter.cpp:34-105).
//
// The reference picks the format by suffix and hands the pixels to vendored
// encoders (miniexr for half-float EXR, jo_jpeg at quality 100 for 8-bit
// JPEG); its EXR path writes the converted half pixels instead of the
// encoded file.  Written here from the file format specifications:
//  * .exr : OpenEXR scanline file, uncompressed, HALF channels B, G, R;
//  * .jpg : baseline JPEG (ITU-T T.81), JFIF YCbCr, no chroma subsampling,
//           quality 100 as the reference asks (every quantiser step 1);
//  * .pfm : portable float map, float32 RGB, bottom row first;
//  * .ppm : binary PPM, 8-bit RGB;
//  * .png : 8-bit RGB, zlib stream of stored (uncompressed) deflate blocks.
// 8-bit values use the reference's conversion (ImageWriter::to_byte,
// src/io/imagewriter.h:48): clamp to [0, 1], gamma 1/2.2, round.
#include <algorithm>
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

// ---- baseline JPEG (ITU-T T.81) ----

// Example quantisation tables of T.81 Annex K.1 / K.2, row-major.
const uint8_t kQuantLuma[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                                14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                                18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                                49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const uint8_t kQuantChroma[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                                  24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                                  99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                                  99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};

// Zig-zag scan: position k of the scan -> row-major index (T.81 Figure 5).
const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Typical Huffman tables of T.81 Annex K.3: code counts per length 1..16, then symbols.
const uint8_t kDcLumaBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const uint8_t kDcChromaBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcLumaBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const uint8_t kAcLumaVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
    0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
    0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
    0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
    0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
    0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
const uint8_t kAcChromaBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
const uint8_t kAcChromaVals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22,
    0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1,
    0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36,
    0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58,
    0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a,
    0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
    0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

// Canonical code table of a DHT (T.81 Annex C): code and length per symbol.
struct Huff {
    uint16_t code[256] = {};
    uint8_t len[256] = {};
    Huff(const uint8_t bits[16], const uint8_t* vals) {
        uint16_t c = 0;
        size_t k = 0;
        for (int l = 1; l <= 16; l++) {
            for (int i = 0; i < bits[l - 1]; i++, k++) {
                code[vals[k]] = c++;
                len[vals[k]] = (uint8_t)l;
            }
            c <<= 1;
        }
    }
};

struct BitWriter {
    std::string& out;
    uint32_t acc = 0;
    int n = 0;
    explicit BitWriter(std::string& o) : out(o) {}
    void put(uint32_t bits, int len) {
        for (int i = len - 1; i >= 0; i--) {
            acc = (acc << 1) | ((bits >> i) & 1u);
            if (++n == 8) {
                out += (char)acc;
                if (acc == 0xFF) out += '\0';  // byte stuffing
                acc = 0;
                n = 0;
            }
        }
    }
    void flush() {
        while (n) put(1, 1);  // pad with 1 bits
    }
};

// Magnitude category of a coefficient and its appended bits (T.81 F.1.2.1).
int category(int v) {
    int a = v < 0 ? -v : v, c = 0;
    while (a) {
        c++;
        a >>= 1;
    }
    return c;
}
uint32_t magnitude_bits(int v, int cat) { return v >= 0 ? (uint32_t)v : (uint32_t)(v + (1 << cat) - 1); }

void jpeg_segment(std::string& s, uint8_t marker, const std::string& body) {
    s += (char)0xFF;
    s += (char)marker;
    const size_t n = body.size() + 2;
    s += (char)(n >> 8);
    s += (char)(n & 0xFF);
    s += body;
}

std::string encode_jpeg(const float* rgb, uint32_t w, uint32_t h, int quality) {
    // quantiser steps: the Annex K tables scaled by the usual quality rule
    // (quality 50 = the tables; 100 = every step 1)
    const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
    uint8_t q[2][64];
    for (int i = 0; i < 64; i++) {
        const int a = (kQuantLuma[i] * scale + 50) / 100, b = (kQuantChroma[i] * scale + 50) / 100;
        q[0][i] = (uint8_t)std::min(255, std::max(1, a));
        q[1][i] = (uint8_t)std::min(255, std::max(1, b));
    }
    const Huff dc[2] = {Huff(kDcLumaBits, kDcVals), Huff(kDcChromaBits, kDcVals)};
    const Huff ac[2] = {Huff(kAcLumaBits, kAcLumaVals), Huff(kAcChromaBits, kAcChromaVals)};

    std::string s("\xFF\xD8", 2);  // SOI
    jpeg_segment(s, 0xE0, std::string("JFIF\0\x01\x01\0\0\x01\0\x01\0\0", 14));
    for (int t = 0; t < 2; t++) {
        std::string b(1, (char)t);  // 8-bit table t
        for (int k = 0; k < 64; k++) b += (char)q[t][kZigzag[k]];
        jpeg_segment(s, 0xDB, b);
    }
    {
        std::string b;
        b += (char)8;
        b += (char)(h >> 8);
        b += (char)(h & 0xFF);
        b += (char)(w >> 8);
        b += (char)(w & 0xFF);
        b += (char)3;
        for (int c = 0; c < 3; c++) {
            b += (char)(c + 1);
            b += (char)0x11;          // no subsampling
            b += (char)(c ? 1 : 0);   // quantisation table
        }
        jpeg_segment(s, 0xC0, b);     // SOF0: baseline
    }
    const struct {
        uint8_t cls_id;
        const uint8_t* bits;
        const uint8_t* vals;
    } dht[4] = {{0x00, kDcLumaBits, kDcVals}, {0x10, kAcLumaBits, kAcLumaVals},
                {0x01, kDcChromaBits, kDcVals}, {0x11, kAcChromaBits, kAcChromaVals}};
    for (const auto& t : dht) {
        std::string b(1, (char)t.cls_id);
        int total = 0;
        for (int l = 0; l < 16; l++) {
            b += (char)t.bits[l];
            total += t.bits[l];
        }
        b.append(reinterpret_cast<const char*>(t.vals), (size_t)total);
        jpeg_segment(s, 0xC4, b);
    }
    jpeg_segment(s, 0xDA, std::string("\x03\x01\x00\x02\x11\x03\x11\x00\x3F\x00", 10));  // SOS

    // DCT basis: cos((2x + 1) u pi / 16), with C(0) = 1/sqrt(2)
    double basis[8][8];
    for (int u = 0; u < 8; u++)
        for (int x = 0; x < 8; x++)
            basis[u][x] = std::cos((2 * x + 1) * u * M_PI / 16.0) * (u ? 1.0 : M_SQRT1_2);
    BitWriter bw(s);
    int pred[3] = {0, 0, 0};
    for (uint32_t by = 0; by < h; by += 8) {
        for (uint32_t bx = 0; bx < w; bx += 8) {
            double blk[3][64];
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 8; x++) {
                    const uint32_t sx = std::min(bx + x, w - 1), sy = std::min(by + y, h - 1);  // edge replicate
                    const float* p = rgb + ((size_t)sy * w + sx) * 3;
                    const double r = to_byte(p[0]), g = to_byte(p[1]), b = to_byte(p[2]);
                    blk[0][y * 8 + x] = 0.299 * r + 0.587 * g + 0.114 * b - 128.0;
                    blk[1][y * 8 + x] = -0.168736 * r - 0.331264 * g + 0.5 * b;
                    blk[2][y * 8 + x] = 0.5 * r - 0.418688 * g - 0.081312 * b;
                }
            for (int c = 0; c < 3; c++) {
                const int t = c ? 1 : 0;
                double tmp[64];
                for (int y = 0; y < 8; y++)  // rows
                    for (int u = 0; u < 8; u++) {
                        double a = 0;
                        for (int x = 0; x < 8; x++) a += blk[c][y * 8 + x] * basis[u][x];
                        tmp[y * 8 + u] = a;
                    }
                int coef[64];
                for (int v = 0; v < 8; v++)  // columns
                    for (int u = 0; u < 8; u++) {
                        double a = 0;
                        for (int y = 0; y < 8; y++) a += tmp[y * 8 + u] * basis[v][y];
                        const int qv = (int)std::lround(a / 4.0 / q[t][v * 8 + u]);
                        coef[v * 8 + u] = std::max(-1023, std::min(1023, qv));
                    }
                // DC difference
                const int diff = coef[0] - pred[c];
                pred[c] = coef[0];
                const int dcat = category(diff);
                bw.put(dc[t].code[dcat], dc[t].len[dcat]);
                if (dcat) bw.put(magnitude_bits(diff, dcat), dcat);
                // AC run lengths in zig-zag order
                int run = 0;
                for (int k = 1; k < 64; k++) {
                    const int v = coef[kZigzag[k]];
                    if (v == 0) {
                        run++;
                        continue;
                    }
                    while (run > 15) {
                        bw.put(ac[t].code[0xF0], ac[t].len[0xF0]);  // ZRL
                        run -= 16;
                    }
                    const int cat = category(v);
                    const int sym = (run << 4) | cat;
                    bw.put(ac[t].code[sym], ac[t].len[sym]);
                    bw.put(magnitude_bits(v, cat), cat);
                    run = 0;
                }
                if (run) bw.put(ac[t].code[0x00], ac[t].len[0x00]);  // EOB
            }
        }
    }
    bw.flush();
    s += "\xFF\xD9";  // EOI
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
        } else if (ends_with(p, ".jpg") || ends_with(p, ".jpeg")) {
            if (w > 65535 || h > 65535) return fail(RB_ERR_INVALID, "JPEG images are at most 65535 x 65535");
            data = encode_jpeg(rgb, w, h, 100);
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
            return fail(RB_ERR_UNSUPPORTED, "%s: unknown image suffix (exr, jpg, png, ppm, pfm)", path);
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
