// Baseline JPEG (ITU-T T.81, sequential DCT, Huffman) of an 8-bit grayscale
// picture: the format the reference's vorticity pictures are written in
// (ComputeUnitOutput.hpp:181-188 hands its picture to a JPEG encoder with
// quality 100).  Written from the standard: level shift by 128, an 8x8
// forward DCT in double precision, quantisation by the luminance table of
// Annex K.1 scaled by the quality, zig-zag order, and the typical Huffman
// tables of Annex K.3 (DC and AC luminance).  Blocks past the picture's edge
// repeat its last row and column.  The output is a JFIF file.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace tdlb {
namespace jpeg {

// T.81 Figure A.6: position in the 8x8 block of the n-th zig-zag coefficient
constexpr uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
// Annex K.1, Table K.1: luminance quantisation table (row-major)
constexpr uint8_t kLumaQuant[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                                    14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                                    18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                                    49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
// Annex K.3, Tables K.3 and K.5: code counts per length 1..16 and symbols
constexpr uint8_t kDcBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
constexpr uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
constexpr uint8_t kAcBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
constexpr uint8_t kAcVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
    0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
    0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
    0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
    0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
    0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

struct Code {
    uint16_t bits = 0;
    uint8_t len = 0;
};

// canonical codes of a table given as counts per length (T.81 Annex C)
inline void build_codes(const uint8_t* counts, const uint8_t* vals, Code* out) {
    uint16_t code = 0;
    int n = 0;
    for (int len = 1; len <= 16; len++) {
        for (int c = 0; c < counts[len - 1]; c++, n++) out[vals[n]] = Code{code++, (uint8_t)len};
        code <<= 1;
    }
}

class BitWriter {
public:
    explicit BitWriter(std::vector<uint8_t>& out) : out_(out) {}
    void put(uint32_t bits, int len) {
        for (int b = len - 1; b >= 0; b--) {
            acc_ = (uint8_t)((acc_ << 1) | ((bits >> b) & 1u));
            if (++n_ == 8) flush_byte();
        }
    }
    void finish() {  // pad the last byte with 1-bits (T.81 F.1.2.3)
        while (n_ != 0) put(1, 1);
    }

private:
    void flush_byte() {
        out_.push_back(acc_);
        if (acc_ == 0xff) out_.push_back(0x00);  // byte stuffing
        acc_ = 0;
        n_ = 0;
    }
    std::vector<uint8_t>& out_;
    uint8_t acc_ = 0;
    int n_ = 0;
};

// magnitude category and the value's low bits (T.81 F.1.2.1)
inline int category(int v, uint32_t& bits) {
    const int a = v < 0 ? -v : v;
    int n = 0;
    while ((a >> n) != 0) n++;
    bits = v < 0 ? (uint32_t)(v + (1 << n) - 1) : (uint32_t)v;
    return n;
}

inline void put16(std::vector<uint8_t>& o, int v) {
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)(v & 0xff));
}

// The JPEG file of a width x height picture, pixel (x, y) at pict[y * width + x].
inline std::vector<uint8_t> encode_gray(const uint8_t* pict, int width, int height, int quality) {
    std::vector<uint8_t> o;
    if (width <= 0 || height <= 0 || width > 65535 || height > 65535) return o;
    quality = quality < 1 ? 1 : quality > 100 ? 100 : quality;
    const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
    uint8_t q[64];
    for (int n = 0; n < 64; n++) {
        int v = (kLumaQuant[n] * scale + 50) / 100;
        q[n] = (uint8_t)(v < 1 ? 1 : v > 255 ? 255 : v);
    }
    Code dc[12], ac[256];
    build_codes(kDcBits, kDcVals, dc);
    build_codes(kAcBits, kAcVals, ac);

    // SOI, JFIF APP0
    const uint8_t head[] = {0xff, 0xd8, 0xff, 0xe0, 0x00, 0x10, 'J', 'F', 'I', 'F', 0x00,
                            0x01, 0x01, 0x00, 0x00, 0x01, 0x00, 0x01, 0x00, 0x00};
    o.insert(o.end(), head, head + sizeof head);
    // DQT: table 0, 8-bit entries in zig-zag order
    o.push_back(0xff); o.push_back(0xdb); put16(o, 67); o.push_back(0x00);
    for (int n = 0; n < 64; n++) o.push_back(q[kZigzag[n]]);
    // SOF0: 8-bit samples, one component (id 1, 1x1 sampling, quantisation table 0)
    o.push_back(0xff); o.push_back(0xc0); put16(o, 11); o.push_back(8);
    put16(o, height); put16(o, width);
    o.push_back(1); o.push_back(1); o.push_back(0x11); o.push_back(0);
    // DHT: DC table 0, AC table 0
    o.push_back(0xff); o.push_back(0xc4); put16(o, 2 + 1 + 16 + 12); o.push_back(0x00);
    o.insert(o.end(), kDcBits, kDcBits + 16); o.insert(o.end(), kDcVals, kDcVals + 12);
    o.push_back(0xff); o.push_back(0xc4); put16(o, 2 + 1 + 16 + 162); o.push_back(0x10);
    o.insert(o.end(), kAcBits, kAcBits + 16); o.insert(o.end(), kAcVals, kAcVals + 162);
    // SOS: component 1 with tables 0/0, spectral selection 0..63
    o.push_back(0xff); o.push_back(0xda); put16(o, 8); o.push_back(1);
    o.push_back(1); o.push_back(0x00); o.push_back(0); o.push_back(63); o.push_back(0);

    double cosv[8][8];
    for (int x = 0; x < 8; x++)
        for (int u = 0; u < 8; u++) cosv[x][u] = std::cos((2 * x + 1) * u * M_PI / 16.0);
    BitWriter bw(o);
    int prev_dc = 0;
    for (int by = 0; by < height; by += 8)
        for (int bx = 0; bx < width; bx += 8) {
            double f[8][8];
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 8; x++) {
                    const int px = bx + x < width ? bx + x : width - 1;
                    const int py = by + y < height ? by + y : height - 1;
                    f[y][x] = (double)pict[(size_t)py * width + px] - 128.0;
                }
            int coef[64];
            for (int v = 0; v < 8; v++)
                for (int u = 0; u < 8; u++) {
                    double s = 0.0;
                    for (int y = 0; y < 8; y++)
                        for (int x = 0; x < 8; x++) s += f[y][x] * cosv[x][u] * cosv[y][v];
                    const double cu = u == 0 ? M_SQRT1_2 : 1.0, cv = v == 0 ? M_SQRT1_2 : 1.0;
                    const double F = 0.25 * cu * cv * s;
                    long r = std::lround(F / q[v * 8 + u]);
                    const long lim = (u == 0 && v == 0) ? 2047 : 1023;
                    coef[v * 8 + u] = (int)(r < -lim ? -lim : r > lim ? lim : r);
                }
            // DC difference
            uint32_t bits = 0;
            const int diff = coef[0] - prev_dc;
            prev_dc = coef[0];
            int n = category(diff, bits);
            bw.put(dc[n].bits, dc[n].len);
            if (n) bw.put(bits, n);
            // AC run lengths in zig-zag order
            int run = 0;
            for (int z = 1; z < 64; z++) {
                const int c = coef[kZigzag[z]];
                if (c == 0) {
                    run++;
                    continue;
                }
                while (run > 15) {
                    bw.put(ac[0xf0].bits, ac[0xf0].len);  // ZRL: 16 zeros
                    run -= 16;
                }
                n = category(c, bits);
                const int sym = (run << 4) | n;
                bw.put(ac[sym].bits, ac[sym].len);
                bw.put(bits, n);
                run = 0;
            }
            if (run > 0) bw.put(ac[0x00].bits, ac[0x00].len);  // EOB
        }
    bw.finish();
    o.push_back(0xff); o.push_back(0xd9);  // EOI
    return o;
}

inline bool write_gray(const std::string& path, const uint8_t* pict, int width, int height, int quality) {
    const std::vector<uint8_t> data = encode_gray(pict, width, height, quality);
    if (data.empty()) return false;
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const bool ok = std::fwrite(data.data(), 1, data.size(), fp) == data.size();
    return std::fclose(fp) == 0 && ok;
}

}  // namespace jpeg
}  // namespace tdlb
