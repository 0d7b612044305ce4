// exr_decode.cpp - OpenEXR scanline images to RGBA32F for sky images.
//
// The reference reads .exr skies through the `image` crate (Cargo.toml:
// image 0.25 with the "exr" feature; world.rs:1872-1873); its own 4k sky
// (assets/skybox/kloppenheim_06_puresky_4k.exr) is a 32-bit float RGBA image
// with PIZ compression. This decoder follows the OpenEXR file layout
// (magic, attribute header, line offset table, chunks of (y, size, data))
// and its compression methods for single-part scanline files:
//   NONE, ZIPS (1 line) and ZIP (16 lines): zlib, then the byte predictor
//     (d[i] = d[i-1] + d[i] - 128) and the split of even/odd bytes;
//   PIZ (32 lines): a bitmap of the 16-bit values in use and its reverse
//     lookup table, canonical Huffman coding of 16-bit words with 6-bit code
//     lengths (run-lengths for zero lengths) and a run symbol, and a 2D Haar
//     wavelet per channel plane (14-bit or modular 16-bit lifting);
//   a chunk whose stored size equals its raw size is stored uncompressed.
// HALF, FLOAT and UINT channels; R, G, B and A are taken by name (a missing A
// is 1, missing colour channels are 0). Values are returned unchanged
// (linear), as the crate's into_rgba32f returns them.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "image_io.h"

namespace ptcore {

namespace {

bool fail(std::string& err, const std::string& m) {
    err = "OpenEXR: " + m;
    return false;
}

uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
uint16_t rd16(const uint8_t* p) { return (uint16_t)(p[0] | p[1] << 8); }

float half_to_float(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 31u, m = h & 1023u;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) bits = s;
        else {  // subnormal half: exact in float
            float f = std::ldexp((float)m, -24);
            std::memcpy(&bits, &f, 4);
            bits |= s;
        }
    } else if (e == 31) {
        bits = s | 0x7f800000u | (m << 13);
    } else {
        bits = s | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// ---------------------------------------------------------------- PIZ: Huffman
constexpr int HUF_ENCSIZE = (1 << 16) + 1;
constexpr int HUF_DECBITS = 14;
constexpr int HUF_DECSIZE = 1 << HUF_DECBITS;
constexpr int SHORT_ZEROCODE_RUN = 59, LONG_ZEROCODE_RUN = 63, SHORTEST_LONG_RUN = 2 + LONG_ZEROCODE_RUN - SHORT_ZEROCODE_RUN;

struct BitReader {
    const uint8_t* p;
    const uint8_t* end;
    uint64_t c = 0;
    int lc = 0;
    bool get(int n, uint64_t& v) {
        while (lc < n) {
            if (p >= end) return false;
            c = (c << 8) | *p++;
            lc += 8;
        }
        lc -= n;
        v = (c >> lc) & ((1ull << n) - 1);
        return true;
    }
};

struct HufDec {
    int len = 0;             // short code length, 0 when none
    int lit = 0;             // short code symbol, or the count of long codes
    std::vector<int> longs;  // symbols of long codes with this 14-bit prefix
};

bool huf_uncompress(const uint8_t* in, size_t n_in, uint16_t* out, size_t n_out, std::string& err) {
    if (n_in == 0) return n_out == 0 ? true : fail(err, "empty Huffman block");
    if (n_in < 20) return fail(err, "short Huffman header");
    const uint32_t im = rd32(in), iM = rd32(in + 4), n_bits = rd32(in + 12);
    if (im >= (uint32_t)HUF_ENCSIZE || iM >= (uint32_t)HUF_ENCSIZE || im > iM) return fail(err, "bad Huffman table range");
    // code lengths, 6 bits each, with run-lengths of zero lengths
    std::vector<uint64_t> hcode(HUF_ENCSIZE, 0);
    BitReader br{in + 20, in + n_in};
    for (uint32_t i = im; i <= iM; ++i) {
        uint64_t l;
        if (!br.get(6, l)) return fail(err, "truncated Huffman table");
        if (l == (uint64_t)LONG_ZEROCODE_RUN) {
            uint64_t run;
            if (!br.get(8, run)) return fail(err, "truncated Huffman table");
            run += SHORTEST_LONG_RUN;
            if (i + run > iM + 1) return fail(err, "Huffman zero run past the table");
            for (uint64_t k = 0; k < run; ++k) hcode[i + k] = 0;
            i += (uint32_t)run - 1;
        } else if (l >= (uint64_t)SHORT_ZEROCODE_RUN) {
            const uint32_t run = (uint32_t)l - SHORT_ZEROCODE_RUN + 2;
            if (i + run > iM + 1) return fail(err, "Huffman zero run past the table");
            for (uint32_t k = 0; k < run; ++k) hcode[i + k] = 0;
            i += run - 1;
        } else {
            hcode[i] = l;
        }
    }
    const uint8_t* bits = br.p;  // the code stream starts at the next whole byte
    if ((uint64_t)n_bits > 8ull * (size_t)(in + n_in - bits)) return fail(err, "Huffman bit count past the block");
    // canonical codes: longest codes first, counting up within a length
    {
        uint64_t n[59] = {0};
        for (int i = 0; i < HUF_ENCSIZE; ++i) n[hcode[i]] += 1;
        uint64_t c = 0;
        for (int i = 58; i > 0; --i) {
            const uint64_t nc = (c + n[i]) >> 1;
            n[i] = c;
            c = nc;
        }
        for (int i = 0; i < HUF_ENCSIZE; ++i) {
            const uint64_t l = hcode[i];
            if (l > 0) hcode[i] = l | (n[l]++ << 6);
        }
    }
    // decoding table: 14-bit prefixes; longer codes listed per prefix
    std::vector<HufDec> dec(HUF_DECSIZE);
    for (uint32_t i = im; i <= iM; ++i) {
        const uint64_t code = hcode[i] >> 6;
        const int l = (int)(hcode[i] & 63);
        if (l == 0) continue;
        if (code >> l) return fail(err, "bad Huffman code");
        if (l > HUF_DECBITS) {
            HufDec& d = dec[code >> (l - HUF_DECBITS)];
            if (d.len) return fail(err, "Huffman prefix clash");
            d.lit++;
            d.longs.push_back((int)i);
        } else {
            const uint64_t first = code << (HUF_DECBITS - l), count = 1ull << (HUF_DECBITS - l);
            for (uint64_t k = 0; k < count; ++k) {
                HufDec& d = dec[first + k];
                if (d.len || !d.longs.empty()) return fail(err, "Huffman prefix clash");
                d.len = l;
                d.lit = (int)i;
            }
        }
    }
    // decode; symbol iM is the run code: the next 8 bits repeat the previous word
    const int rlc = (int)iM;
    size_t o = 0;
    uint64_t scanned = 0;
    const uint64_t scan_budget = 64ull * n_out + (1ull << 20);
    uint64_t c = 0;
    int lc = 0;
    const uint8_t* p = bits;
    const uint8_t* ie = bits + ((size_t)n_bits + 7) / 8;
    auto emit = [&](int sym) -> bool {
        if (sym == rlc) {
            if (lc < 8) {
                if (p >= ie) return false;
                c = (c << 8) | *p++;
                lc += 8;
            }
            lc -= 8;
            const uint32_t cs = (uint32_t)((c >> lc) & 255u);
            if (o + cs > n_out || o == 0) return false;
            const uint16_t s = out[o - 1];
            for (uint32_t k = 0; k < cs; ++k) out[o++] = s;
            return true;
        }
        if (o >= n_out) return false;
        out[o++] = (uint16_t)sym;
        return true;
    };
    while (p < ie) {
        c = (c << 8) | *p++;
        lc += 8;
        while (lc >= HUF_DECBITS) {
            const HufDec& d = dec[(c >> (lc - HUF_DECBITS)) & (HUF_DECSIZE - 1)];
            if (d.len) {
                lc -= d.len;
                if (!emit(d.lit)) return fail(err, "Huffman data overrun");
            } else {
                if (d.longs.empty()) return fail(err, "invalid Huffman code");
                // codes longer than the table are found by a scan of their prefix's list; a table that
                // makes those scans long (many long codes under one prefix) is refused instead of decoded
                scanned += d.longs.size();
                if (scanned > scan_budget) return fail(err, "Huffman table too costly to decode");
                bool found = false;
                for (int sym : d.longs) {
                    const int l = (int)(hcode[sym] & 63);
                    while (lc < l && p < ie) {
                        c = (c << 8) | *p++;
                        lc += 8;
                    }
                    if (lc >= l && (hcode[sym] >> 6) == ((c >> (lc - l)) & ((1ull << l) - 1))) {
                        lc -= l;
                        if (!emit(sym)) return fail(err, "Huffman data overrun");
                        found = true;
                        break;
                    }
                }
                if (!found) return fail(err, "invalid long Huffman code");
            }
        }
    }
    // the remaining short codes in the last bits
    const int pad = (8 - (int)(n_bits & 7)) & 7;
    c >>= pad;
    lc -= pad;
    while (lc > 0) {
        const HufDec& d = dec[(c << (HUF_DECBITS - lc)) & (HUF_DECSIZE - 1)];
        if (!d.len || d.len > lc) return fail(err, "invalid trailing Huffman code");
        lc -= d.len;
        if (!emit(d.lit)) return fail(err, "Huffman data overrun");
    }
    if (o != n_out) return fail(err, "Huffman data shorter than the block");
    return true;
}

// ---------------------------------------------------------------- PIZ: wavelet
inline void wdec14(uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
    const int16_t ls = (int16_t)l, hs = (int16_t)h;
    const int hi = hs;
    const int ai = ls + (hi & 1) + (hi >> 1);
    a = (uint16_t)(int16_t)ai;
    b = (uint16_t)(int16_t)(ai - hi);
}
inline void wdec16(uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
    const int m = l, d = h;
    const int bb = (m - (d >> 1)) & 0xffff;
    const int aa = (d + bb - 0x8000) & 0xffff;
    b = (uint16_t)bb;
    a = (uint16_t)aa;
}

void wav2_decode(uint16_t* in, int nx, int ox, int ny, int oy, uint16_t mx) {
    const bool w14 = mx < (1 << 14);
    const int n = nx > ny ? ny : nx;
    int p = 1;
    while (p <= n) p <<= 1;
    p >>= 1;
    int p2 = p;
    p >>= 1;
    auto dec = [&](uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
        if (w14) wdec14(l, h, a, b);
        else wdec16(l, h, a, b);
    };
    while (p >= 1) {
        uint16_t* py = in;
        uint16_t* ey = in + (ptrdiff_t)oy * (ny - p2);
        const int oy1 = oy * p, oy2 = oy * p2, ox1 = ox * p, ox2 = ox * p2;
        uint16_t i00, i01, i10, i11;
        for (; py <= ey; py += oy2) {
            uint16_t* px = py;
            uint16_t* ex = py + (ptrdiff_t)ox * (nx - p2);
            for (; px <= ex; px += ox2) {
                uint16_t* p01 = px + ox1;
                uint16_t* p10 = px + oy1;
                uint16_t* p11 = p10 + ox1;
                dec(*px, *p10, i00, i10);
                dec(*p01, *p11, i01, i11);
                dec(i00, i01, *px, *p01);
                dec(i10, i11, *p10, *p11);
            }
            if (nx & p) {  // odd column
                uint16_t* p10 = px + oy1;
                dec(*px, *p10, i00, *p10);
                *px = i00;
            }
        }
        if (ny & p) {  // odd line
            uint16_t* px = py;
            uint16_t* ex = py + (ptrdiff_t)ox * (nx - p2);
            for (; px <= ex; px += ox2) {
                uint16_t* p01 = px + ox1;
                dec(*px, *p01, i00, *p01);
                *px = i00;
            }
        }
        p2 = p;
        p >>= 1;
    }
}

struct Channel {
    std::string name;
    int type;  // 0 UINT, 1 HALF, 2 FLOAT
    int bytes() const { return type == 1 ? 2 : 4; }
};

// PIZ chunk of `lines` lines, `w` wide, into scanline-interleaved bytes
bool piz_chunk(const uint8_t* in, size_t n, const std::vector<Channel>& chans, int w, int lines, std::vector<uint8_t>& out,
               std::string& err) {
    size_t words = 0;
    for (const Channel& ch : chans) words += (size_t)w * lines * (ch.bytes() / 2);
    if (n < 4) return fail(err, "short PIZ chunk");
    const uint16_t min_nz = rd16(in), max_nz = rd16(in + 2);
    size_t pos = 4;
    if (max_nz >= 8192) return fail(err, "bad PIZ bitmap range");
    std::vector<uint8_t> bitmap(8192, 0);
    if (min_nz <= max_nz) {
        const size_t len = (size_t)max_nz - min_nz + 1;
        if (pos + len > n) return fail(err, "truncated PIZ bitmap");
        std::memcpy(&bitmap[min_nz], in + pos, len);
        pos += len;
    }
    std::vector<uint16_t> lut(65536, 0);
    int k = 0;
    for (int i = 0; i < 65536; ++i)
        if (i == 0 || (bitmap[i >> 3] & (1 << (i & 7)))) lut[k++] = (uint16_t)i;
    const uint16_t max_value = (uint16_t)(k - 1);
    if (pos + 4 > n) return fail(err, "truncated PIZ chunk");
    const uint32_t len = rd32(in + pos);
    pos += 4;
    if (len > n - pos) return fail(err, "PIZ Huffman data past the chunk");
    std::vector<uint16_t> tmp(words);
    if (!huf_uncompress(in + pos, len, tmp.data(), words, err)) return false;
    size_t start = 0;
    std::vector<size_t> starts;
    for (const Channel& ch : chans) {
        const int size = ch.bytes() / 2;
        starts.push_back(start);
        for (int j = 0; j < size; ++j) wav2_decode(&tmp[start + j], w, size, lines, w * size, max_value);
        start += (size_t)w * lines * size;
    }
    for (uint16_t& v : tmp) v = lut[v];
    out.resize(words * 2);
    size_t o = 0;
    for (int y = 0; y < lines; ++y)
        for (size_t c = 0; c < chans.size(); ++c) {
            const size_t cnt = (size_t)w * (chans[c].bytes() / 2);
            const uint16_t* src = &tmp[starts[c] + (size_t)y * cnt];
            for (size_t i = 0; i < cnt; ++i) {
                out[o++] = (uint8_t)(src[i] & 255);
                out[o++] = (uint8_t)(src[i] >> 8);
            }
        }
    return true;
}

bool zip_chunk(const uint8_t* in, size_t n, size_t raw, std::vector<uint8_t>& out, std::string& err) {
    std::vector<uint8_t> t(raw);
    uLongf len = (uLongf)raw;
    if (uncompress(t.data(), &len, in, (uLong)n) != Z_OK || len != raw) return fail(err, "corrupt ZIP chunk");
    for (size_t i = 1; i < raw; ++i) t[i] = (uint8_t)(t[i - 1] + t[i] - 128);
    out.resize(raw);
    const size_t half = (raw + 1) / 2;
    for (size_t i = 0; i < raw; ++i) out[i] = (i & 1) ? t[half + i / 2] : t[i / 2];
    return true;
}

}  // namespace

bool decode_exr(const uint8_t* data, size_t size, FloatImage& out, std::string& err) {
    if (size < 8 || rd32(data) != 20000630u) return fail(err, "not an OpenEXR file");
    const uint32_t version = rd32(data + 4);
    if ((version & 255) != 2) return fail(err, "unsupported version");
    if (version & 0x200) return fail(err, "tiled images are not supported");
    if (version & 0x1800) return fail(err, "multi-part or deep images are not supported");
    size_t pos = 8;
    std::vector<Channel> chans;
    int compression = -1, line_order = 0;
    int32_t dw[4] = {0, 0, -1, -1};
    bool have_dw = false;
    auto cstr = [&](std::string& s) {
        s.clear();
        while (pos < size && data[pos]) s.push_back((char)data[pos++]);
        if (pos >= size) return false;
        ++pos;
        return true;
    };
    for (;;) {
        std::string name, type;
        if (!cstr(name)) return fail(err, "truncated header");
        if (name.empty()) break;
        if (!cstr(type) || pos + 4 > size) return fail(err, "truncated header");
        const uint32_t len = rd32(data + pos);
        pos += 4;
        if (len > size - pos) return fail(err, "truncated header attribute");
        const uint8_t* v = data + pos;
        if (name == "channels" && type == "chlist") {
            size_t q = 0;
            while (q < len && v[q]) {
                Channel ch;
                while (q < len && v[q]) ch.name.push_back((char)v[q++]);
                ++q;
                if (q + 16 > len) return fail(err, "bad channel list");
                ch.type = (int)rd32(v + q);
                const int32_t xs = (int32_t)rd32(v + q + 8), ys = (int32_t)rd32(v + q + 12);
                if (ch.type < 0 || ch.type > 2) return fail(err, "bad channel type");
                if (xs != 1 || ys != 1) return fail(err, "subsampled channels are not supported");
                chans.push_back(ch);
                q += 16;
            }
        } else if (name == "compression" && len >= 1) {
            compression = v[0];
        } else if (name == "dataWindow" && len >= 16) {
            for (int k = 0; k < 4; ++k) dw[k] = (int32_t)rd32(v + 4 * k);
            have_dw = true;
        } else if (name == "lineOrder" && len >= 1) {
            line_order = v[0];
        }
        pos += len;
    }
    (void)line_order;  // chunks carry their own first line
    if (chans.empty() || !have_dw) return fail(err, "header without channels or data window");
    const int64_t w = (int64_t)dw[2] - dw[0] + 1, h = (int64_t)dw[3] - dw[1] + 1;
    if (w <= 0 || h <= 0 || w > 65536 || h > 65536) return fail(err, "bad data window");
    if ((uint64_t)w * (uint64_t)h > MAX_IMAGE_PIXELS)
        return fail(err, "image of " + std::to_string(w) + "x" + std::to_string(h) + " exceeds the image size limit");
    int lines_per_chunk;
    switch (compression) {
        case 0: case 2: lines_per_chunk = 1; break;   // NONE, ZIPS
        case 3: lines_per_chunk = 16; break;          // ZIP
        case 4: lines_per_chunk = 32; break;          // PIZ
        default: return fail(err, "compression method " + std::to_string(compression) + " is not supported");
    }
    size_t line_bytes = 0;
    for (const Channel& ch : chans) line_bytes += (size_t)w * ch.bytes();
    const size_t nchunks = (size_t)((h + lines_per_chunk - 1) / lines_per_chunk);
    if (pos + nchunks * 8 > size) return fail(err, "truncated offset table");
    std::vector<uint8_t> pixels((size_t)h * line_bytes);
    std::vector<uint8_t> buf;
    for (size_t c = 0; c < nchunks; ++c) {
        const uint64_t off = (uint64_t)rd32(data + pos + 8 * c) | (uint64_t)rd32(data + pos + 8 * c + 4) << 32;
        if (off + 8 > size) return fail(err, "chunk offset past the file");
        const int32_t y = (int32_t)rd32(data + off);
        const uint32_t csize = rd32(data + off + 4);
        if (csize > size - off - 8) return fail(err, "chunk past the file");
        const int64_t line0 = (int64_t)y - dw[1];
        if (line0 < 0 || line0 >= h || line0 % lines_per_chunk) return fail(err, "bad chunk line");
        const int lines = (int)std::min<int64_t>(lines_per_chunk, h - line0);
        const size_t raw = (size_t)lines * line_bytes;
        const uint8_t* cdata = data + off + 8;
        uint8_t* dst = &pixels[(size_t)line0 * line_bytes];
        if (csize == raw || compression == 0) {
            if (csize != raw) return fail(err, "uncompressed chunk of the wrong size");
            std::memcpy(dst, cdata, raw);
        } else if (compression == 4) {
            if (!piz_chunk(cdata, csize, chans, (int)w, lines, buf, err)) return false;
            std::memcpy(dst, buf.data(), raw);
        } else {
            if (!zip_chunk(cdata, csize, raw, buf, err)) return false;
            std::memcpy(dst, buf.data(), raw);
        }
    }
    // channels by name into RGBA
    int idx[4] = {-1, -1, -1, -1};
    size_t offs[16];
    size_t o = 0;
    for (size_t c = 0; c < chans.size(); ++c) {
        const std::string& nm = chans[c].name;
        if (nm == "R") idx[0] = (int)c;
        else if (nm == "G") idx[1] = (int)c;
        else if (nm == "B") idx[2] = (int)c;
        else if (nm == "A") idx[3] = (int)c;
        if (c < 16) offs[c] = o;
        o += (size_t)w * chans[c].bytes();
    }
    if (chans.size() > 16) return fail(err, "too many channels");
    if (idx[0] < 0 && idx[1] < 0 && idx[2] < 0) return fail(err, "no R, G or B channel");
    out.width = (uint32_t)w;
    out.height = (uint32_t)h;
    out.rgba.assign((size_t)w * h * 4, 0.0f);
    for (int64_t yy = 0; yy < h; ++yy) {
        const uint8_t* line = &pixels[(size_t)yy * line_bytes];
        for (int64_t x = 0; x < w; ++x) {
            float* px = &out.rgba[((size_t)yy * w + x) * 4];
            for (int k = 0; k < 4; ++k) {
                if (idx[k] < 0) {
                    px[k] = k == 3 ? 1.0f : 0.0f;
                    continue;
                }
                const Channel& ch = chans[idx[k]];
                const uint8_t* s = line + offs[idx[k]] + (size_t)x * ch.bytes();
                if (ch.type == 1) px[k] = half_to_float(rd16(s));
                else if (ch.type == 2) {
                    uint32_t b = rd32(s);
                    std::memcpy(&px[k], &b, 4);
                } else {
                    px[k] = (float)rd32(s);
                }
            }
        }
    }
    return true;
}

}  // namespace ptcore
