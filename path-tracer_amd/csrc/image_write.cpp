// image_write.cpp - encoders for the output image of `pt-render render`:
// PNG, baseline JPEG and lossless WebP, from RGB8 rows; and, for scene
// export, RGBA PNG textures, uncompressed FLOAT OpenEXR skies and exact
// Radiance RGBE skies (run-length coded scanlines, refused unless every value
// decodes back unchanged).
//
// The reference saves its tone-mapped RGB8 image with image::save_buffer
// (headless.rs:163-170) and picks the format from the file extension: .png,
// .jpg/.jpeg or .webp (headless.rs:118-127). The `image` crate (Cargo.toml:
// image 0.25) is not part of the reference tree. These encoders follow the
// format specifications:
//   PNG   ISO/IEC 15948 (zlib stream, per-row filter chosen by the minimum sum
//         of absolute differences);
//   JPEG  ITU-T T.81 baseline sequential DCT with a JFIF APP0 segment; quality
//         75 by default (the crate's JpegEncoder default), the T.81 Annex K.1
//         example quantisation tables scaled by the IJG quality formula, every
//         component at full resolution, Huffman tables generated for the image
//         by the Annex K.2 procedure;
//   WebP  lossless (VP8L, RFC 9649): subtract-green and predictor transforms,
//         then LZ77 with a hash chain and one group of prefix codes.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <queue>
#include <string>
#include <vector>

#include "../../include/pt_assets.h"
#include "pt_internal.h"

namespace ptcore {

namespace {

void put_be32(std::vector<uint8_t>& o, uint32_t v) {
    for (int k = 3; k >= 0; --k) o.push_back((uint8_t)(v >> (8 * k)));
}
void put_be16(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}
void put_le32(std::vector<uint8_t>& o, uint32_t v) {
    for (int k = 0; k < 4; ++k) o.push_back((uint8_t)(v >> (8 * k)));
}

// ---------------------------------------------------------------- PNG

int paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

}  // namespace

// PNG of 8-bit RGB (channels 3) or RGBA (channels 4) rows
bool encode_png_channels(const uint8_t* rgb, size_t stride, uint32_t w, uint32_t h, int channels,
                         std::vector<uint8_t>& out, std::string& err) {
    const int C = channels == 4 ? 4 : 3;
    const size_t row = (size_t)w * C;
    std::vector<uint8_t> raw((row + 1) * h), cand(row);
    std::vector<uint8_t> zero(row, 0);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* cur = rgb + y * stride;
        const uint8_t* prev = y ? rgb + (y - 1) * stride : zero.data();
        uint8_t* dst = &raw[y * (row + 1)];
        uint64_t best = UINT64_MAX;
        for (int f = 0; f < 5; ++f) {
            uint64_t cost = 0;
            for (size_t i = 0; i < row; ++i) {
                int a = i >= (size_t)C ? cur[i - C] : 0, b = prev[i], c = i >= (size_t)C ? prev[i - C] : 0;
                int pred = f == 0 ? 0 : f == 1 ? a : f == 2 ? b : f == 3 ? (a + b) >> 1 : paeth(a, b, c);
                uint8_t v = (uint8_t)(cur[i] - pred);
                cand[i] = v;
                cost += v < 128 ? v : 256 - v;
            }
            if (cost < best) {
                best = cost;
                dst[0] = (uint8_t)f;
                std::memcpy(dst + 1, cand.data(), row);
            }
        }
    }
    uLongf zlen = compressBound(raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), raw.size(), 6) != Z_OK) {
        err = "PNG: zlib compression failed";
        return false;
    }
    z.resize(zlen);
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    out.assign(sig, sig + 8);
    auto chunk = [&](const char* type, const uint8_t* data, size_t n) {
        put_be32(out, (uint32_t)n);
        size_t at = out.size();
        out.insert(out.end(), type, type + 4);
        if (n) out.insert(out.end(), data, data + n);
        put_be32(out, (uint32_t)crc32(0, out.data() + at, (uInt)(n + 4)));
    };
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    // 8-bit truecolour (with alpha for 4 channels), deflate, adaptive filters, no interlace
    ihdr.insert(ihdr.end(), {8, (uint8_t)(C == 4 ? 6 : 2), 0, 0, 0});
    chunk("IHDR", ihdr.data(), ihdr.size());
    chunk("IDAT", z.data(), z.size());
    chunk("IEND", nullptr, 0);
    return true;
}

// Single-part scanline OpenEXR, uncompressed, FLOAT channels A, B, G, R (the
// channel list is sorted by name), one scanline per block.
bool encode_exr_float(const float* rgba, uint32_t w, uint32_t h, std::vector<uint8_t>& out, std::string& err) {
    if (!w || !h) {
        err = "EXR: empty image";
        return false;
    }
    out.clear();
    auto u32 = [&](uint32_t v) { put_le32(out, v); };
    auto str = [&](const char* t) { out.insert(out.end(), t, t + std::strlen(t) + 1); };
    auto attr = [&](const char* name, const char* type, uint32_t size) {
        str(name);
        str(type);
        u32(size);
    };
    auto f32 = [&](float f) {
        uint32_t b;
        std::memcpy(&b, &f, 4);
        u32(b);
    };
    u32(20000630u);  // magic
    u32(2u);         // version 2, scanline, single part
    attr("channels", "chlist", 4 * (2 + 16) + 1);
    for (const char* ch : {"A", "B", "G", "R"}) {
        str(ch);
        u32(2u);  // FLOAT
        out.insert(out.end(), {0, 0, 0, 0});  // pLinear, reserved
        u32(1u);
        u32(1u);  // x and y sampling
    }
    out.push_back(0);
    attr("compression", "compression", 1);
    out.push_back(0);  // NO_COMPRESSION
    attr("dataWindow", "box2i", 16);
    u32(0), u32(0), u32(w - 1), u32(h - 1);
    attr("displayWindow", "box2i", 16);
    u32(0), u32(0), u32(w - 1), u32(h - 1);
    attr("lineOrder", "lineOrder", 1);
    out.push_back(0);  // INCREASING_Y
    attr("pixelAspectRatio", "float", 4);
    f32(1.0f);
    attr("screenWindowCenter", "v2f", 8);
    f32(0.0f), f32(0.0f);
    attr("screenWindowWidth", "float", 4);
    f32(1.0f);
    out.push_back(0);  // end of header
    const size_t table = out.size();
    out.resize(table + (size_t)h * 8);
    const uint32_t line_bytes = w * 4 * 4;
    for (uint32_t y = 0; y < h; ++y) {
        const uint64_t at = out.size();
        std::memcpy(&out[table + (size_t)y * 8], &at, 8);
        u32(y);
        u32(line_bytes);
        for (int c : {3, 2, 1, 0})  // A, B, G, R
            for (uint32_t x = 0; x < w; ++x) f32(rgba[((size_t)y * w + x) * 4 + c]);
    }
    return true;
}

static bool hdr_fail(std::string& err, const char* m) {
    err = std::string("HDR: ") + m;
    return false;
}

bool encode_hdr_exact(const float* rgba, uint32_t w, uint32_t h, std::vector<uint8_t>& out, std::string& err) {
    if (w == 0 || h == 0 || w > 65535 || h > 65535) return hdr_fail(err, "bad HDR size");
    char head[96];
    std::snprintf(head, sizeof(head), "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", h, w);
    out.assign(head, head + std::strlen(head));
    std::vector<uint8_t> scan((size_t)w * 4);
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            const float* p = &rgba[((size_t)y * w + x) * 4];
            uint8_t* e = &scan[(size_t)x * 4];
            const float mx = std::fmax(p[0], std::fmax(p[1], p[2]));
            e[0] = e[1] = e[2] = e[3] = 0;
            if (mx > 0.0f) {
                int ex;
                std::frexp(mx, &ex);  // mx = f * 2^ex, f in [0.5, 1)
                if (ex + 128 < 1 || ex + 128 > 255) return hdr_fail(err, "sky value outside the RGBE exponent range");
                for (int c = 0; c < 3; ++c) e[c] = (uint8_t)(int)std::ldexp(p[c], 8 - ex);
                e[3] = (uint8_t)(ex + 128);
            }
            // exact only if the decoder (c * 2^(E - 136)) gives every value back
            for (int c = 0; c < 3; ++c) {
                const float back = e[3] ? std::ldexp((float)e[c], (int)e[3] - 136) : 0.0f;
                if (!(back == p[c]) || std::signbit(p[c]))
                    return hdr_fail(err, "sky is not exactly representable in RGBE (write it as .exr)");
            }
            if (p[3] != 1.0f) return hdr_fail(err, "sky alpha is not 1 (RGBE holds no alpha; write it as .exr)");
        }
        const bool rle = w >= 8 && w < 32768;
        if (!rle) {
            out.insert(out.end(), scan.begin(), scan.end());
            continue;
        }
        out.push_back(2);
        out.push_back(2);
        out.push_back((uint8_t)(w >> 8));
        out.push_back((uint8_t)(w & 0xFF));
        for (int c = 0; c < 4; ++c) {  // each component run-length coded on its own
            uint32_t x = 0;
            while (x < w) {
                uint32_t run = 1;
                while (x + run < w && run < 127 && scan[(size_t)(x + run) * 4 + c] == scan[(size_t)x * 4 + c]) ++run;
                if (run >= 3) {
                    out.push_back((uint8_t)(128 + run));
                    out.push_back(scan[(size_t)x * 4 + c]);
                    x += run;
                    continue;
                }
                // literal span up to the next run of 3 or 128 values
                uint32_t n = 0;
                while (x + n < w && n < 128) {
                    const uint32_t i = x + n;
                    if (i + 2 < w && scan[(size_t)i * 4 + c] == scan[(size_t)(i + 1) * 4 + c] &&
                        scan[(size_t)i * 4 + c] == scan[(size_t)(i + 2) * 4 + c])
                        break;
                    ++n;
                }
                out.push_back((uint8_t)n);
                for (uint32_t k = 0; k < n; ++k) out.push_back(scan[(size_t)(x + k) * 4 + c]);
                x += n;
            }
        }
    }
    return true;
}

namespace {

bool encode_png(const uint8_t* rgb, size_t stride, uint32_t w, uint32_t h, std::vector<uint8_t>& out,
                std::string& err) {
    return encode_png_channels(rgb, stride, w, h, 3, out, err);
}

// ---------------------------------------------------------------- JPEG

// T.81 Annex K.1, Tables K.1 (luminance) and K.2 (chrominance), row-major.
const uint8_t kLumaQ[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                            14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                            18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                            49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const uint8_t kChromaQ[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                              24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                              99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                              99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};

// zigzag[k] = row-major index of the k-th coefficient in scan order: the
// anti-diagonals row + col = s, walked upwards (row falling) for even s.
struct Zigzag {
    int zz[64];
    Zigzag() {
        int k = 0;
        for (int s = 0; s < 15; ++s) {
            int lo = std::max(0, s - 7), hi = std::min(s, 7);
            if (s % 2 == 0)
                for (int r = hi; r >= lo; --r) zz[k++] = r * 8 + (s - r);
            else
                for (int r = lo; r <= hi; ++r) zz[k++] = r * 8 + (s - r);
        }
    }
};

void scale_quant(const uint8_t* base, int quality, uint8_t* q) {
    quality = std::min(100, std::max(1, quality));
    int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
    for (int i = 0; i < 64; ++i) q[i] = (uint8_t)std::min(255, std::max(1, (base[i] * scale + 50) / 100));
}

// Huffman table specification (BITS, HUFFVAL) for the symbol counts `freq`
// (256 symbols) by T.81 Annex K.2: code sizes by repeated merging of the two
// least frequent entries (ties to the higher symbol), with a reserved symbol
// of count 1 so no code is all ones; lengths above 16 folded back (Figure
// K.3); symbols listed by code size, then symbol value (Figure K.4).
struct HuffSpec {
    uint8_t bits[17] = {};
    std::vector<uint8_t> vals;
    uint16_t code[256] = {};
    uint8_t size[256] = {};
};

void build_huffman(const uint32_t* counts, HuffSpec& hs) {
    int64_t freq[257];
    int codesize[257], others[257];
    for (int i = 0; i < 256; ++i) freq[i] = counts[i];
    freq[256] = 1;
    for (int i = 0; i < 257; ++i) codesize[i] = 0, others[i] = -1;
    for (;;) {
        int v1 = -1, v2 = -1;
        for (int i = 0; i < 257; ++i)
            if (freq[i] && (v1 < 0 || freq[i] <= freq[v1])) v1 = i;
        for (int i = 0; i < 257; ++i)
            if (freq[i] && i != v1 && (v2 < 0 || freq[i] <= freq[v2])) v2 = i;
        if (v2 < 0) break;
        freq[v1] += freq[v2];
        freq[v2] = 0;
        ++codesize[v1];
        while (others[v1] >= 0) v1 = others[v1], ++codesize[v1];
        others[v1] = v2;
        ++codesize[v2];
        while (others[v2] >= 0) v2 = others[v2], ++codesize[v2];
    }
    int bits[64] = {};
    for (int i = 0; i < 257; ++i)
        if (codesize[i]) ++bits[codesize[i]];
    for (int i = 63; i > 16; --i)
        while (bits[i] > 0) {
            int j = i - 2;
            while (bits[j] == 0) --j;
            bits[i] -= 2;
            bits[i - 1] += 1;
            bits[j + 1] += 2;
            bits[j] -= 1;
        }
    int i = 16;
    while (bits[i] == 0) --i;
    bits[i] -= 1;  // drop the reserved symbol's code (the last, longest one)
    hs.vals.clear();
    for (int len = 1; len < 64; ++len)
        for (int s = 0; s < 256; ++s)
            if (codesize[s] == len) hs.vals.push_back((uint8_t)s);
    int n = 0;
    for (int len = 1; len <= 16; ++len) hs.bits[len] = (uint8_t)bits[len], n += bits[len];
    hs.vals.resize(n);
    // canonical codes (Annex C)
    uint32_t c = 0;
    size_t k = 0;
    for (int len = 1; len <= 16; ++len) {
        for (int m = 0; m < hs.bits[len]; ++m, ++k) {
            hs.code[hs.vals[k]] = (uint16_t)c++;
            hs.size[hs.vals[k]] = (uint8_t)len;
        }
        c <<= 1;
    }
}

struct JpegBits {
    std::vector<uint8_t>& out;
    uint32_t acc = 0;
    int n = 0;
    explicit JpegBits(std::vector<uint8_t>& o) : out(o) {}
    void put(uint32_t v, int len) {
        for (int b = len - 1; b >= 0; --b) {
            acc = acc << 1 | ((v >> b) & 1u);
            if (++n == 8) {
                out.push_back((uint8_t)acc);
                if (acc == 0xff) out.push_back(0);
                acc = 0;
                n = 0;
            }
        }
    }
    void flush() {
        if (n) put((1u << (8 - n)) - 1, 8 - n);
    }
};

int magnitude_category(int v) {
    v = std::abs(v);
    int s = 0;
    while (v) ++s, v >>= 1;
    return s;
}

bool encode_jpeg(const uint8_t* rgb, size_t stride, uint32_t w, uint32_t h, int quality, std::vector<uint8_t>& out,
                 std::string& err) {
    if (w > 65535 || h > 65535) {
        err = "JPEG: image larger than 65535 pixels on a side";
        return false;
    }
    static const Zigzag zig;
    uint8_t q[2][64];
    scale_quant(kLumaQ, quality, q[0]);
    scale_quant(kChromaQ, quality, q[1]);
    // cos((2x+1) u pi / 16) * C(u) / 2
    double basis[8][8];
    for (int u = 0; u < 8; ++u)
        for (int x = 0; x < 8; ++x)
            basis[u][x] = std::cos((2 * x + 1) * u * M_PI / 16.0) * (u ? 0.5 : 0.5 / std::sqrt(2.0));
    const uint32_t bw = (w + 7) / 8, bh = (h + 7) / 8;
    const size_t nblocks = (size_t)bw * bh;
    std::vector<int16_t> coef(nblocks * 3 * 64);  // per block: Y, Cb, Cr, zigzag order
    for (uint32_t by = 0; by < bh; ++by)
        for (uint32_t bx = 0; bx < bw; ++bx) {
            double plane[3][64];
            for (int y = 0; y < 8; ++y)
                for (int x = 0; x < 8; ++x) {
                    uint32_t sx = std::min(bx * 8 + x, w - 1), sy = std::min(by * 8 + y, h - 1);
                    const uint8_t* p = rgb + sy * stride + (size_t)sx * 3;
                    double r = p[0], g = p[1], b = p[2];
                    plane[0][y * 8 + x] = 0.299 * r + 0.587 * g + 0.114 * b - 128.0;
                    plane[1][y * 8 + x] = -0.168736 * r - 0.331264 * g + 0.5 * b;
                    plane[2][y * 8 + x] = 0.5 * r - 0.418688 * g - 0.081312 * b;
                }
            for (int c = 0; c < 3; ++c) {
                double tmp[64], f[64];
                for (int y = 0; y < 8; ++y)  // rows: horizontal frequencies
                    for (int u = 0; u < 8; ++u) {
                        double s = 0;
                        for (int x = 0; x < 8; ++x) s += basis[u][x] * plane[c][y * 8 + x];
                        tmp[y * 8 + u] = s;
                    }
                for (int u = 0; u < 8; ++u)  // columns: vertical frequencies
                    for (int v = 0; v < 8; ++v) {
                        double s = 0;
                        for (int y = 0; y < 8; ++y) s += basis[v][y] * tmp[y * 8 + u];
                        f[v * 8 + u] = s;
                    }
                int16_t* dst = &coef[((size_t)by * bw + bx) * 192 + c * 64];
                const uint8_t* qt = q[c ? 1 : 0];
                for (int k = 0; k < 64; ++k) {
                    int idx = zig.zz[k];
                    dst[k] = (int16_t)std::lround(f[idx] / qt[idx]);
                }
            }
        }
    // statistics: table 0 for Y, table 1 for Cb and Cr
    uint32_t dc_count[2][256] = {}, ac_count[2][256] = {};
    auto walk = [&](auto&& emit_dc, auto&& emit_ac) {
        int pred[3] = {0, 0, 0};
        for (size_t b = 0; b < nblocks; ++b)
            for (int c = 0; c < 3; ++c) {
                const int16_t* blk = &coef[b * 192 + c * 64];
                int t = c ? 1 : 0;
                int diff = blk[0] - pred[c];
                pred[c] = blk[0];
                emit_dc(t, diff);
                int run = 0;
                for (int k = 1; k < 64; ++k) {
                    if (blk[k] == 0) {
                        ++run;
                        continue;
                    }
                    while (run > 15) emit_ac(t, 0xf0, 0), run -= 16;
                    emit_ac(t, run << 4 | magnitude_category(blk[k]), blk[k]);
                    run = 0;
                }
                if (run) emit_ac(t, 0x00, 0);
            }
    };
    walk([&](int t, int diff) { ++dc_count[t][magnitude_category(diff)]; },
         [&](int t, int rs, int) { ++ac_count[t][rs]; });
    HuffSpec dc[2], ac[2];
    for (int t = 0; t < 2; ++t) build_huffman(dc_count[t], dc[t]), build_huffman(ac_count[t], ac[t]);

    out.clear();
    out.insert(out.end(), {0xff, 0xd8});                                        // SOI
    out.insert(out.end(), {0xff, 0xe0, 0, 16, 'J', 'F', 'I', 'F', 0, 1, 2, 0, 0, 1, 0, 1, 0, 0});  // APP0 JFIF 1.02
    for (int t = 0; t < 2; ++t) {                                               // DQT
        out.insert(out.end(), {0xff, 0xdb, 0, 67, (uint8_t)t});
        for (int k = 0; k < 64; ++k) out.push_back(q[t][zig.zz[k]]);
    }
    out.insert(out.end(), {0xff, 0xc0, 0, 17, 8});                              // SOF0
    put_be16(out, h);
    put_be16(out, w);
    out.insert(out.end(), {3, 1, 0x11, 0, 2, 0x11, 1, 3, 0x11, 1});
    for (int t = 0; t < 2; ++t)                                                 // DHT
        for (int cls = 0; cls < 2; ++cls) {
            const HuffSpec& hs = cls ? ac[t] : dc[t];
            out.insert(out.end(), {0xff, 0xc4});
            put_be16(out, (uint32_t)(2 + 1 + 16 + hs.vals.size()));
            out.push_back((uint8_t)(cls << 4 | t));
            out.insert(out.end(), hs.bits + 1, hs.bits + 17);
            out.insert(out.end(), hs.vals.begin(), hs.vals.end());
        }
    out.insert(out.end(), {0xff, 0xda, 0, 12, 3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0});  // SOS
    JpegBits bits(out);
    auto extra = [&](int v, int s) {
        if (s) bits.put((uint32_t)(v < 0 ? v + (1 << s) - 1 : v), s);
    };
    walk(
        [&](int t, int diff) {
            int s = magnitude_category(diff);
            bits.put(dc[t].code[s], dc[t].size[s]);
            extra(diff, s);
        },
        [&](int t, int rs, int v) {
            bits.put(ac[t].code[rs], ac[t].size[rs]);
            extra(v, rs & 15);
        });
    bits.flush();
    out.insert(out.end(), {0xff, 0xd9});  // EOI
    return true;
}

// ---------------------------------------------------------------- WebP lossless

struct LsbBits {
    std::vector<uint8_t> buf;
    uint64_t acc = 0;
    int n = 0;
    void put(uint32_t v, int nb) {
        acc |= (uint64_t)v << n;
        n += nb;
        while (n >= 8) {
            buf.push_back((uint8_t)acc);
            acc >>= 8;
            n -= 8;
        }
    }
    void flush() {
        if (n) buf.push_back((uint8_t)acc);
        acc = 0;
        n = 0;
    }
};

// Code lengths (<= limit) of a Huffman code for `freq`. Counts are raised to
// a growing floor until the tree fits the limit; a single used symbol gets
// length 1 (the format codes it with zero bits).
std::vector<uint8_t> limited_lengths(const std::vector<uint32_t>& freq, int limit) {
    const int n = (int)freq.size();
    std::vector<uint8_t> len(n, 0);
    std::vector<int> used;
    for (int i = 0; i < n; ++i)
        if (freq[i]) used.push_back(i);
    if (used.size() == 1) len[used[0]] = 1;
    if (used.size() <= 1) return len;
    for (uint64_t floor_count = 1;; floor_count *= 2) {
        struct Node {
            uint64_t w;
            int id;
        };
        auto cmp = [](const Node& a, const Node& b) { return a.w != b.w ? a.w > b.w : a.id > b.id; };
        std::priority_queue<Node, std::vector<Node>, decltype(cmp)> heap(cmp);
        std::vector<int> parent(2 * used.size(), -1);
        for (size_t i = 0; i < used.size(); ++i) heap.push({std::max<uint64_t>(freq[used[i]], floor_count), (int)i});
        int next = (int)used.size();
        while (heap.size() > 1) {
            Node a = heap.top();
            heap.pop();
            Node b = heap.top();
            heap.pop();
            parent[a.id] = parent[b.id] = next;
            heap.push({a.w + b.w, next++});
        }
        int depth_max = 0;
        for (size_t i = 0; i < used.size(); ++i) {
            int d = 0;
            for (int p = parent[i]; p >= 0; p = parent[p]) ++d;
            len[used[i]] = (uint8_t)d;
            depth_max = std::max(depth_max, d);
        }
        if (depth_max <= limit) return len;
    }
}

// Canonical codes (shortest first, then by symbol), bit-reversed for the
// LSB-first stream. A code with one used symbol is written with zero bits.
struct Prefix {
    std::vector<uint16_t> code;
    std::vector<uint8_t> len;
    void build(const std::vector<uint8_t>& lengths) {
        const int n = (int)lengths.size();
        code.assign(n, 0);
        len = lengths;
        int used = 0;
        for (uint8_t l : lengths) used += l != 0;
        if (used == 1) {
            for (auto& l : len) l = 0;
            return;
        }
        int count[16] = {}, next[16] = {};
        for (uint8_t l : lengths)
            if (l) ++count[l];
        int c = 0;
        for (int l = 1; l < 16; ++l) c = (c + count[l - 1]) << 1, next[l] = c;
        for (int s = 0; s < n; ++s)
            if (lengths[s]) {
                uint32_t v = (uint32_t)next[lengths[s]]++, r = 0;
                for (int b = 0; b < lengths[s]; ++b) r |= ((v >> b) & 1u) << (lengths[s] - 1 - b);
                code[s] = (uint16_t)r;
            }
    }
    void put(LsbBits& bw, int s) const {
        if (len[s]) bw.put(code[s], len[s]);
    }
};

// Writes the code lengths of one prefix code (simple form for at most two
// symbols below 256, otherwise the normal form with run-length tokens 16, 17
// and 18) and returns the code for the data.
Prefix write_prefix_code(LsbBits& bw, const std::vector<uint32_t>& freq) {
    std::vector<int> used;
    for (int i = 0; i < (int)freq.size(); ++i)
        if (freq[i]) used.push_back(i);
    Prefix p;
    if (used.size() <= 2 && (used.empty() || used.back() < 256)) {
        std::vector<uint8_t> lengths(freq.size(), 0);
        int s0 = used.empty() ? 0 : used[0];
        bw.put(1, 1);
        bw.put(used.size() == 2 ? 1 : 0, 1);
        if (s0 < 2) {
            bw.put(0, 1);
            bw.put((uint32_t)s0, 1);
        } else {
            bw.put(1, 1);
            bw.put((uint32_t)s0, 8);
        }
        lengths[s0] = 1;
        if (used.size() == 2) bw.put((uint32_t)used[1], 8), lengths[used[1]] = 1;
        p.build(lengths);
        return p;
    }
    std::vector<uint8_t> lengths = limited_lengths(freq, 15);
    // tokens: (symbol 0..18, extra value)
    std::vector<std::pair<int, int>> tok;
    const int n = (int)lengths.size();
    for (int i = 0; i < n;) {
        int v = lengths[i], r = 1;
        while (i + r < n && lengths[i + r] == v) ++r;
        i += r;
        if (v == 0) {
            while (r >= 11) {
                int k = std::min(r, 138);
                tok.push_back({18, k - 11});
                r -= k;
            }
            if (r >= 3) tok.push_back({17, r - 3}), r = 0;
            while (r--) tok.push_back({0, 0});
        } else {
            tok.push_back({v, 0});
            --r;
            while (r >= 3) {
                int k = std::min(r, 6);
                tok.push_back({16, k - 3});
                r -= k;
            }
            while (r--) tok.push_back({v, 0});
        }
    }
    std::vector<uint32_t> tf(19, 0);
    for (auto& t : tok) ++tf[t.first];
    std::vector<uint8_t> tl = limited_lengths(tf, 7);
    static const int order[19] = {17, 18, 0, 1, 2, 3, 4, 5, 16, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
    int count = 4;
    for (int i = 0; i < 19; ++i)
        if (tl[order[i]]) count = std::max(count, i + 1);
    bw.put(0, 1);
    bw.put((uint32_t)(count - 4), 4);
    for (int i = 0; i < count; ++i) bw.put(tl[order[i]], 3);
    bw.put(0, 1);  // every symbol of the alphabet is coded
    Prefix tp;
    tp.build(tl);
    static const int extra_bits[3] = {2, 3, 7};
    for (auto& t : tok) {
        tp.put(bw, t.first);
        if (t.first >= 16) bw.put((uint32_t)t.second, extra_bits[t.first - 16]);
    }
    p.build(lengths);
    return p;
}

// Prefix coding of a length or distance value v >= 1: (symbol, extra bit count, extra bits).
void prefix_value(uint32_t v, int& sym, int& nbits, uint32_t& extra) {
    uint32_t d = v - 1;
    if (d < 4) {
        sym = (int)d, nbits = 0, extra = 0;
        return;
    }
    int hb = 31 - __builtin_clz(d);
    int second = (int)((d >> (hb - 1)) & 1u);
    nbits = hb - 1;
    extra = d & ((1u << nbits) - 1);
    sym = 2 * hb + second;
}

constexpr int kMaxLength = 4096;
constexpr uint32_t kMaxDistance = (1u << 20) - 121;  // distance codes reach 2^20 = distance + 120

// Entropy-coded image: no colour cache, one prefix-code group (no meta codes
// for the main image), greedy LZ77 (hash of two pixels, bounded chain).
void write_entropy_image(LsbBits& bw, const std::vector<uint32_t>& argb, uint32_t width, bool main_image,
                         bool use_lz77) {
    struct Sym {
        uint32_t v;     // pixel, or length for a copy
        uint32_t dist;  // 0 for a literal
    };
    std::vector<Sym> syms;
    syms.reserve(argb.size());
    const size_t n = argb.size();
    if (use_lz77 && n > 2) {
        const int kHashBits = 18, kChain = 24;
        std::vector<int32_t> head(1u << kHashBits, -1), prev(n, -1);
        auto hash = [&](size_t i) {
            uint64_t k = (uint64_t)argb[i] * 0x9E3779B97F4A7C15ull ^ (uint64_t)argb[i + 1] * 0xC2B2AE3D27D4EB4Full;
            return (uint32_t)(k >> (64 - kHashBits));
        };
        auto insert = [&](size_t i) {
            if (i + 1 >= n) return;
            uint32_t hv = hash(i);
            prev[i] = head[hv];
            head[hv] = (int32_t)i;
        };
        for (size_t i = 0; i < n;) {
            size_t best_len = 0, best_dist = 0;
            if (i + 1 < n) {
                int32_t cand = head[hash(i)];
                for (int c = 0; c < kChain && cand >= 0; ++c, cand = prev[cand]) {
                    size_t dist = i - (size_t)cand;
                    if (dist > kMaxDistance) break;
                    size_t l = 0, lmax = std::min<size_t>(kMaxLength, n - i);
                    while (l < lmax && argb[cand + l] == argb[i + l]) ++l;
                    if (l > best_len) best_len = l, best_dist = dist;
                    if (l == lmax) break;
                }
            }
            if (best_len >= 3) {
                syms.push_back({(uint32_t)best_len, (uint32_t)best_dist});
                for (size_t k = 0; k < best_len; ++k) insert(i + k);
                i += best_len;
            } else {
                syms.push_back({argb[i], 0});
                insert(i);
                ++i;
            }
        }
    } else {
        for (uint32_t v : argb) syms.push_back({v, 0});
    }
    // distance value: 2-D codes 1 (the pixel above) and 2 (the pixel to the left), otherwise distance + 120
    auto dist_code = [&](uint32_t d) -> uint32_t { return d == width ? 1u : d == 1 ? 2u : d + 120u; };
    std::vector<uint32_t> fg(256 + 24, 0), fr(256, 0), fb(256, 0), fa(256, 0), fd(40, 0);
    for (const Sym& s : syms) {
        int sym, nb;
        uint32_t ex;
        if (!s.dist) {
            ++fg[(s.v >> 8) & 255], ++fr[(s.v >> 16) & 255], ++fb[s.v & 255], ++fa[s.v >> 24];
        } else {
            prefix_value(s.v, sym, nb, ex);
            ++fg[256 + sym];
            prefix_value(dist_code(s.dist), sym, nb, ex);
            ++fd[sym];
        }
    }
    bw.put(0, 1);                  // no colour cache
    if (main_image) bw.put(0, 1);  // no meta prefix codes
    Prefix pg = write_prefix_code(bw, fg), pr = write_prefix_code(bw, fr), pb = write_prefix_code(bw, fb),
           pa = write_prefix_code(bw, fa), pd = write_prefix_code(bw, fd);
    for (const Sym& s : syms) {
        if (!s.dist) {
            pg.put(bw, (s.v >> 8) & 255);
            pr.put(bw, (s.v >> 16) & 255);
            pb.put(bw, s.v & 255);
            pa.put(bw, s.v >> 24);
        } else {
            int sym, nb;
            uint32_t ex;
            prefix_value(s.v, sym, nb, ex);
            pg.put(bw, 256 + sym);
            if (nb) bw.put(ex, nb);
            prefix_value(dist_code(s.dist), sym, nb, ex);
            pd.put(bw, sym);
            if (nb) bw.put(ex, nb);
        }
    }
}

inline int ch(uint32_t p, int k) { return (int)((p >> (8 * k)) & 255u); }
// per-channel floor((a + b) / 2)
inline uint32_t avg2(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; ++k) r |= (uint32_t)((ch(a, k) + ch(b, k)) >> 1) << (8 * k);
    return r;
}
inline uint32_t clamp_add_sub_full(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r = 0;
    for (int k = 0; k < 4; ++k) r |= (uint32_t)std::min(255, std::max(0, ch(a, k) + ch(b, k) - ch(c, k))) << (8 * k);
    return r;
}
inline uint32_t clamp_add_sub_half(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; ++k) r |= (uint32_t)std::min(255, std::max(0, ch(a, k) + (ch(a, k) - ch(b, k)) / 2)) << (8 * k);
    return r;
}
inline uint32_t select_pred(uint32_t L, uint32_t T, uint32_t TL) {
    int pl = 0, pt = 0;
    for (int k = 0; k < 4; ++k) {
        int est = ch(L, k) + ch(T, k) - ch(TL, k);
        pl += std::abs(est - ch(L, k));
        pt += std::abs(est - ch(T, k));
    }
    return pl < pt ? L : T;
}

uint32_t predict(int mode, uint32_t L, uint32_t T, uint32_t TR, uint32_t TL) {
    switch (mode) {
        case 0: return 0xff000000u;
        case 1: return L;
        case 2: return T;
        case 3: return TR;
        case 4: return TL;
        case 5: return avg2(avg2(L, TR), T);
        case 6: return avg2(L, TL);
        case 7: return avg2(L, T);
        case 8: return avg2(TL, T);
        case 9: return avg2(T, TR);
        case 10: return avg2(avg2(L, TL), avg2(T, TR));
        case 11: return select_pred(L, T, TL);
        case 12: return clamp_add_sub_full(L, T, TL);
        default: return clamp_add_sub_half(avg2(L, T), TL);
    }
}

// per-channel (a - b) mod 256
inline uint32_t sub_pixels(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; ++k) r |= (uint32_t)((ch(a, k) - ch(b, k)) & 255) << (8 * k);
    return r;
}

bool encode_webp(const uint8_t* rgb, size_t stride, uint32_t w, uint32_t h, std::vector<uint8_t>& out,
                 std::string& err) {
    if (w > 16384 || h > 16384) {
        err = "WebP: image larger than 16384 pixels on a side";
        return false;
    }
    // ARGB with green subtracted from red and blue
    std::vector<uint32_t> px((size_t)w * h);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const uint8_t* p = rgb + y * stride + (size_t)x * 3;
            uint32_t r = (uint8_t)(p[0] - p[1]), g = p[1], b = (uint8_t)(p[2] - p[1]);
            px[(size_t)y * w + x] = 0xff000000u | r << 16 | g << 8 | b;
        }
    const int size_bits = 4;
    const uint32_t bs = 1u << size_bits, tw = (w + bs - 1) / bs, th = (h + bs - 1) / bs;
    auto neighbours_pred = [&](int mode, uint32_t x, uint32_t y) -> uint32_t {
        if (y == 0) return x == 0 ? 0xff000000u : px[x - 1];
        if (x == 0) return px[(size_t)(y - 1) * w];
        size_t i = (size_t)y * w + x;
        uint32_t TR = px[i - w + 1];  // on the last column this is the first pixel of the current row
        return predict(mode, px[i - 1], px[i - w], TR, px[i - w - 1]);
    };
    std::vector<uint32_t> modes((size_t)tw * th), res((size_t)w * h);
    for (uint32_t ty = 0; ty < th; ++ty)
        for (uint32_t tx = 0; tx < tw; ++tx) {
            uint64_t best = UINT64_MAX;
            int best_mode = 1;
            for (int m = 0; m < 14; ++m) {
                uint64_t cost = 0;
                for (uint32_t y = ty * bs; y < std::min(h, (ty + 1) * bs); ++y)
                    for (uint32_t x = tx * bs; x < std::min(w, (tx + 1) * bs); ++x) {
                        uint32_t r = sub_pixels(px[(size_t)y * w + x], neighbours_pred(m, x, y));
                        for (int k = 0; k < 3; ++k) cost += (uint64_t)std::abs((int)(int8_t)ch(r, k));
                    }
                if (cost < best) best = cost, best_mode = m;
            }
            modes[(size_t)ty * tw + tx] = (uint32_t)best_mode << 8;
            for (uint32_t y = ty * bs; y < std::min(h, (ty + 1) * bs); ++y)
                for (uint32_t x = tx * bs; x < std::min(w, (tx + 1) * bs); ++x)
                    res[(size_t)y * w + x] = sub_pixels(px[(size_t)y * w + x], neighbours_pred(best_mode, x, y));
        }
    LsbBits bw;
    bw.put(0x2f, 8);
    bw.put(w - 1, 14);
    bw.put(h - 1, 14);
    bw.put(0, 1);  // alpha unused
    bw.put(0, 3);  // version 0
    bw.put(1, 1), bw.put(2, 2);  // subtract-green transform
    bw.put(1, 1), bw.put(0, 2), bw.put(size_bits - 2, 3);  // predictor transform
    write_entropy_image(bw, modes, tw, false, false);
    bw.put(0, 1);  // no more transforms
    write_entropy_image(bw, res, w, true, true);
    bw.flush();
    const uint32_t payload = (uint32_t)bw.buf.size();
    out.clear();
    out.insert(out.end(), {'R', 'I', 'F', 'F'});
    put_le32(out, 4 + 8 + payload + (payload & 1));
    out.insert(out.end(), {'W', 'E', 'B', 'P', 'V', 'P', '8', 'L'});
    put_le32(out, payload);
    out.insert(out.end(), bw.buf.begin(), bw.buf.end());
    if (payload & 1) out.push_back(0);
    return true;
}

bool encode_image(int format, const uint8_t* rgb, size_t stride, uint32_t w, uint32_t h, int quality,
                  std::vector<uint8_t>& out, std::string& err) {
    switch (format) {
        case PT_IMAGE_PNG: return encode_png(rgb, stride, w, h, out, err);
        case PT_IMAGE_JPEG: return encode_jpeg(rgb, stride, w, h, quality ? quality : 75, out, err);
        case PT_IMAGE_WEBP: return encode_webp(rgb, stride, w, h, out, err);
        default: err = "unknown image format"; return false;
    }
}

}  // namespace

}  // namespace ptcore

using namespace ptcore;

extern "C" {

pt_status pt_image_format_from_path(const char* path, pt_image_format* format) {
    if (!path || !format) return set_error(PT_ERR_INVALID, "pt_image_format_from_path: null argument");
    std::string p(path);
    size_t dot = p.find_last_of('.'), slash = p.find_last_of('/');
    std::string ext = dot == std::string::npos || (slash != std::string::npos && dot < slash) ? "" : p.substr(dot + 1);
    for (char& c : ext) c = (char)std::tolower((unsigned char)c);
    if (ext == "png") *format = PT_IMAGE_PNG;
    else if (ext == "jpg" || ext == "jpeg") *format = PT_IMAGE_JPEG;
    else if (ext == "webp") *format = PT_IMAGE_WEBP;
    else return set_error(PT_ERR_INVALID, "'%s' must have a .png, .jpg, .jpeg, or .webp extension", path);
    return PT_OK;
}

pt_status pt_image_encode(pt_image_format format, const uint8_t* rgb, size_t row_stride, uint32_t width,
                          uint32_t height, uint32_t quality, uint8_t** out, size_t* size) {
    if (!rgb || !out || !size) return set_error(PT_ERR_INVALID, "pt_image_encode: null argument");
    *out = nullptr;
    *size = 0;
    if (!width || !height) return set_error(PT_ERR_INVALID, "pt_image_encode: empty image");
    if (quality > 100) return set_error(PT_ERR_INVALID, "pt_image_encode: quality %u above 100", quality);
    if (row_stride == 0) row_stride = (size_t)width * 3;
    if (row_stride < (size_t)width * 3) return set_error(PT_ERR_INVALID, "pt_image_encode: row stride below width * 3");
    try {
        std::vector<uint8_t> bytes;
        std::string err;
        if (!encode_image(format, rgb, row_stride, width, height, (int)quality, bytes, err))
            return set_error(PT_ERR_INVALID, "%s", err.c_str());
        uint8_t* buf = (uint8_t*)std::malloc(bytes.size());
        if (!buf) return set_error(PT_ERR_MEMORY, "out of host memory");
        std::memcpy(buf, bytes.data(), bytes.size());
        *out = buf;
        *size = bytes.size();
    } catch (const std::bad_alloc&) {
        return set_error(PT_ERR_MEMORY, "out of host memory while encoding the image");
    }
    return PT_OK;
}

void pt_image_free(uint8_t* data) { std::free(data); }

pt_status pt_image_save(const char* path, const uint8_t* rgb, uint32_t width, uint32_t height) {
    pt_image_format format;
    pt_status st = pt_image_format_from_path(path, &format);
    if (st != PT_OK) return st;
    uint8_t* bytes = nullptr;
    size_t size = 0;
    st = pt_image_encode(format, rgb, 0, width, height, 0, &bytes, &size);
    if (st != PT_OK) return st;
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        pt_image_free(bytes);
        return set_error(PT_ERR_INVALID, "cannot save %s: cannot open for writing", path);
    }
    bool ok = std::fwrite(bytes, 1, size, f) == size;
    ok = std::fclose(f) == 0 && ok;
    pt_image_free(bytes);
    return ok ? PT_OK : set_error(PT_ERR_INVALID, "cannot save %s: write failed", path);
}

}  // extern "C"
