// jpeg_decode.cpp - JPEG images (glTF textures) to 8-bit grey or RGB samples.
//
// The reference decodes glTF images through the `image` crate (Cargo.toml:
// image 0.25 with the "jpeg" feature, via gltf::import_slice, world.rs:887-889)
// and converts them to RGBA8. The crate is not part of the reference tree;
// this decoder follows ITU-T T.81 (JPEG) and JFIF 1.02:
//   frames   SOF0 (baseline), SOF1 (extended sequential), SOF2 (progressive),
//            Huffman coding, 8-bit samples, 1 or 3 components, any sampling
//            factors from 1 to 4;
//   scans    interleaved and single-component, restart intervals, spectral
//            selection and successive approximation (T.81 G.1.2);
//   output   dequantisation and an exact separable float inverse DCT, samples
//            rounded and clamped to 0..255; chroma upsampled by the triangle
//            filter for 2x1 and 2x2 subsampling of components wider than 2
//            samples (the "fancy" upsampling of the IJG library, which the
//            crate's decoder also applies) and by replication otherwise;
//            YCbCr to RGB by the JFIF equations, no conversion for Adobe RGB
//            (APP14 transform 0) or components named 'R', 'G', 'B'.
// Arithmetic coding, lossless and 12-bit frames, and 4-component (CMYK/YCCK)
// images are rejected with a message.
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
    err = "JPEG: " + m;
    return false;
}

struct Zigzag {
    int zz[80];  // row-major index of scan position k; positions 64.. absorb corrupt run lengths
    Zigzag() {
        int k = 0;
        for (int s = 0; s < 15; ++s) {
            int lo = std::max(0, s - 7), hi = std::min(s, 7);
            if (s % 2 == 0)
                for (int r = hi; r >= lo; --r) zz[k++] = r * 8 + (s - r);
            else
                for (int r = lo; r <= hi; ++r) zz[k++] = r * 8 + (s - r);
        }
        for (; k < 80; ++k) zz[k] = 63;
    }
};

struct Huffman {
    bool defined = false;
    uint8_t vals[256] = {};
    int maxcode[18] = {}, valptr[17] = {}, mincode[17] = {};
    // 9-bit lookup: (length << 8) | value, 0 when the code is longer
    uint16_t fast[512] = {};
    bool build(const uint8_t* bits, const uint8_t* v, int nvals) {
        std::memcpy(vals, v, (size_t)nvals);
        int code = 0, k = 0;
        for (int len = 1; len <= 16; ++len) {
            valptr[len] = k;
            mincode[len] = code;
            code += bits[len - 1];
            k += bits[len - 1];
            maxcode[len] = bits[len - 1] ? code - 1 : -1;
            if (code > (1 << len)) return false;  // over-subscribed
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        std::memset(fast, 0, sizeof(fast));
        code = 0;
        k = 0;
        for (int len = 1; len <= 9; ++len) {
            for (int m = 0; m < bits[len - 1]; ++m, ++k, ++code) {
                const int shift = 9 - len;
                for (int f = 0; f < (1 << shift); ++f) fast[(code << shift) | f] = (uint16_t)((len << 8) | vals[k]);
            }
            code <<= 1;
        }
        defined = true;
        return true;
    }
};

// Entropy-coded segment reader: bytes are consumed MSB first, 0xFF00 is a
// stuffed 0xFF, and a marker ends the data (zeros are fed after it).
struct Bits {
    const uint8_t* p;
    const uint8_t* end;
    uint64_t acc = 0;
    int n = 0;
    bool at_marker = false;
    void fill() {
        while (n <= 56) {
            uint32_t b = 0;
            if (!at_marker && p < end) {
                b = *p;
                if (b == 0xFF) {
                    const uint8_t nx = p + 1 < end ? p[1] : 0xD9;
                    if (nx == 0x00) {
                        p += 2;
                    } else {
                        at_marker = true;
                        b = 0;
                    }
                } else {
                    ++p;
                }
            }
            acc |= (uint64_t)b << (56 - n);
            n += 8;
        }
    }
    uint32_t peek(int k) {
        if (n < k) fill();
        return (uint32_t)(acc >> (64 - k));
    }
    void skip(int k) {
        acc <<= k;
        n -= k;
    }
    uint32_t get(int k) {
        if (k == 0) return 0;
        uint32_t v = peek(k);
        skip(k);
        return v;
    }
    // restart: drop the partial byte and step over the RSTn marker
    bool restart() {
        acc = 0;
        n = 0;
        at_marker = false;
        while (p + 1 < end && !(p[0] == 0xFF && p[1] >= 0xD0 && p[1] <= 0xD7)) ++p;
        if (p + 1 >= end) return false;
        p += 2;
        return true;
    }
};

int decode_symbol(Bits& b, const Huffman& h) {
    const uint32_t look = b.peek(9);
    const uint16_t f = h.fast[look];
    if (f) {
        b.skip(f >> 8);
        return f & 0xFF;
    }
    uint32_t code = b.peek(16);
    for (int len = 10; len <= 16; ++len) {
        const int c = (int)(code >> (16 - len));
        if (c <= h.maxcode[len]) {
            b.skip(len);
            return h.vals[h.valptr[len] + c - h.mincode[len]];
        }
    }
    return -1;
}

inline int extend(uint32_t v, int s) { return s == 0 ? 0 : (v < (1u << (s - 1)) ? (int)v - (1 << s) + 1 : (int)v); }

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int bw = 0, bh = 0;    // blocks per row / column, padded to whole MCUs
    int pw = 0, ph = 0;    // sample width / height of this component
    int td = 0, ta = 0;    // Huffman tables of the current scan
    int dc_pred = 0;
    bool q_latched = false;
    uint16_t q[64] = {};   // quantisation table in row-major order
    std::vector<int16_t> coef;  // bw * bh blocks of 64 row-major coefficients
};

struct Decoder {
    const uint8_t* data;
    size_t size;
    std::string& err;
    Zigzag zig;
    uint16_t qt[4][64] = {};  // row-major
    bool qt_defined[4] = {};
    Huffman dc[4], ac[4];
    std::vector<Component> comp;
    int width = 0, height = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    bool progressive = false, frame = false, adobe = false, jfif = false;
    int adobe_transform = -1;
    int restart_interval = 0;
    int eobrun = 0;

    Decoder(const uint8_t* d, size_t n, std::string& e) : data(d), size(n), err(e) {}

    bool read_dqt(const uint8_t* p, size_t len) {
        size_t o = 0;
        while (o < len) {
            const int pq = p[o] >> 4, tq = p[o] & 15;
            if (tq > 3 || pq > 1) return fail(err, "bad DQT");
            const size_t need = 1 + 64 * (pq ? 2 : 1);
            if (o + need > len) return fail(err, "truncated DQT");
            for (int k = 0; k < 64; ++k)
                qt[tq][zig.zz[k]] = pq ? (uint16_t)(p[o + 1 + 2 * k] << 8 | p[o + 2 + 2 * k]) : p[o + 1 + k];
            qt_defined[tq] = true;
            o += need;
        }
        return true;
    }
    bool read_dht(const uint8_t* p, size_t len) {
        size_t o = 0;
        while (o < len) {
            if (o + 17 > len) return fail(err, "truncated DHT");
            const int tc = p[o] >> 4, th = p[o] & 15;
            if (tc > 1 || th > 3) return fail(err, "bad DHT");
            int nv = 0;
            for (int k = 0; k < 16; ++k) nv += p[o + 1 + k];
            if (nv > 256 || o + 17 + nv > len) return fail(err, "bad DHT size");
            if (!(tc ? ac[th] : dc[th]).build(p + o + 1, p + o + 17, nv)) return fail(err, "bad Huffman table");
            o += 17 + nv;
        }
        return true;
    }
    bool read_sof(const uint8_t* p, size_t len, int marker) {
        if (frame) return fail(err, "more than one frame");
        if (len < 6) return fail(err, "truncated SOF");
        if (p[0] != 8) return fail(err, "only 8-bit samples are supported");
        height = p[1] << 8 | p[2];
        width = p[3] << 8 | p[4];
        const int nc = p[5];
        if (width == 0 || height == 0) return fail(err, "zero image size (DNL is not supported)");
        if ((uint64_t)width * (uint64_t)height > MAX_IMAGE_PIXELS)
            return fail(err, "image of " + std::to_string(width) + "x" + std::to_string(height) + " exceeds the image size limit");
        if (nc != 1 && nc != 3) return fail(err, std::to_string(nc) + "-component images are not supported");
        if (len < 6 + 3 * (size_t)nc) return fail(err, "truncated SOF");
        comp.resize(nc);
        for (int i = 0; i < nc; ++i) {
            Component& c = comp[i];
            c.id = p[6 + 3 * i];
            c.h = p[7 + 3 * i] >> 4;
            c.v = p[7 + 3 * i] & 15;
            c.tq = p[8 + 3 * i];
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) return fail(err, "bad component parameters");
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        mcux = (width + 8 * hmax - 1) / (8 * hmax);
        mcuy = (height + 8 * vmax - 1) / (8 * vmax);
        for (Component& c : comp) {
            if (hmax % c.h || vmax % c.v) return fail(err, "sampling factors that do not divide the maximum");
            c.bw = mcux * c.h;
            c.bh = mcuy * c.v;
            c.pw = (width * c.h + hmax - 1) / hmax;
            c.ph = (height * c.v + vmax - 1) / vmax;
            c.coef.assign((size_t)c.bw * c.bh * 64, 0);
        }
        progressive = marker == 0xC2;
        frame = true;
        return true;
    }

    int16_t* block(Component& c, int bx, int by) { return &c.coef[((size_t)by * c.bw + bx) * 64]; }

    // sequential: one whole block (T.81 F.2.2)
    bool decode_sequential(Bits& b, Component& c, int16_t* blk) {
        const int t = decode_symbol(b, dc[c.td]);
        if (t < 0 || t > 11) return fail(err, "bad DC code");
        c.dc_pred += extend(b.get(t), t);
        blk[0] = (int16_t)c.dc_pred;
        for (int k = 1; k < 64;) {
            const int rs = decode_symbol(b, ac[c.ta]);
            if (rs < 0) return fail(err, "bad AC code");
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r != 15) break;
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) return fail(err, "AC run past the block");
            blk[zig.zz[k]] = (int16_t)extend(b.get(s), s);
            ++k;
        }
        return true;
    }
    // progressive scans (T.81 G.1.2)
    bool decode_dc_first(Bits& b, Component& c, int16_t* blk, int al) {
        const int t = decode_symbol(b, dc[c.td]);
        if (t < 0 || t > 11) return fail(err, "bad DC code");
        c.dc_pred += extend(b.get(t), t);
        blk[0] = (int16_t)(c.dc_pred * (1 << al));
        return true;
    }
    void decode_dc_refine(Bits& b, int16_t* blk, int al) {
        if (b.get(1)) blk[0] = (int16_t)(blk[0] | (1 << al));
    }
    bool decode_ac_first(Bits& b, const Component& c, int16_t* blk, int ss, int se, int al) {
        if (eobrun > 0) {
            --eobrun;
            return true;
        }
        for (int k = ss; k <= se; ++k) {
            const int rs = decode_symbol(b, ac[c.ta]);
            if (rs < 0) return fail(err, "bad AC code");
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r < 15) {
                    eobrun = (1 << r) - 1;
                    if (r) eobrun += (int)b.get(r);
                    break;
                }
                k += 15;
                continue;
            }
            k += r;
            if (k > 63) return fail(err, "AC run past the block");
            blk[zig.zz[k]] = (int16_t)(extend(b.get(s), s) * (1 << al));
        }
        return true;
    }
    bool decode_ac_refine(Bits& b, const Component& c, int16_t* blk, int ss, int se, int al) {
        const int p1 = 1 << al, m1 = -(1 << al);
        auto refine = [&](int16_t& v) {
            if (b.get(1) && (v & p1) == 0) v = (int16_t)(v + (v >= 0 ? p1 : m1));
        };
        int k = ss;
        if (eobrun == 0) {
            for (; k <= se; ++k) {
                const int rs = decode_symbol(b, ac[c.ta]);
                if (rs < 0) return fail(err, "bad AC code");
                int r = rs >> 4;
                const int s = rs & 15;
                int val = 0;
                if (s) {
                    if (s != 1) return fail(err, "bad refinement magnitude");
                    val = b.get(1) ? p1 : m1;
                } else if (r != 15) {
                    eobrun = 1 << r;
                    if (r) eobrun += (int)b.get(r);
                    break;
                }
                // step over r coefficients that are still zero, refining the nonzero ones passed
                while (k <= se) {
                    int16_t& v = blk[zig.zz[k]];
                    if (v != 0) {
                        refine(v);
                    } else {
                        if (r == 0) break;
                        --r;
                    }
                    ++k;
                }
                if (val && k <= se) blk[zig.zz[k]] = (int16_t)val;
            }
        }
        if (eobrun > 0) {
            for (; k <= se; ++k) {
                int16_t& v = blk[zig.zz[k]];
                if (v != 0) refine(v);
            }
            --eobrun;
        }
        return true;
    }

    bool read_sos(const uint8_t* p, size_t len, size_t& pos) {
        if (!frame) return fail(err, "scan before frame");
        if (len < 1) return fail(err, "truncated SOS");
        const int ns = p[0];
        if (ns < 1 || ns > 4 || len < 4 + 2 * (size_t)ns) return fail(err, "bad SOS");
        std::vector<Component*> sc;
        for (int i = 0; i < ns; ++i) {
            const int cid = p[1 + 2 * i];
            Component* c = nullptr;
            for (Component& x : comp)
                if (x.id == cid) c = &x;
            if (!c) return fail(err, "scan names an unknown component");
            c->td = p[2 + 2 * i] >> 4;
            c->ta = p[2 + 2 * i] & 15;
            if (c->td > 3 || c->ta > 3) return fail(err, "bad table selector");
            if (!c->q_latched) {
                if (!qt_defined[c->tq]) return fail(err, "missing quantisation table");
                std::memcpy(c->q, qt[c->tq], sizeof(c->q));
                c->q_latched = true;
            }
            sc.push_back(c);
        }
        const int ss = p[1 + 2 * ns], se = p[2 + 2 * ns], ah = p[3 + 2 * ns] >> 4, al = p[3 + 2 * ns] & 15;
        if (progressive) {
            if (ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && ns != 1) || al > 13)
                return fail(err, "bad progressive scan parameters");
        } else if (ss != 0 || se != 63 || ah != 0 || al != 0) {
            return fail(err, "bad sequential scan parameters");
        }
        const bool dc_scan = progressive && ss == 0;
        for (Component* c : sc) {
            if ((!progressive || (dc_scan && ah == 0)) && !dc[c->td].defined) return fail(err, "missing DC table");
            if ((!progressive || !dc_scan) && !ac[c->ta].defined) return fail(err, "missing AC table");
            c->dc_pred = 0;
        }
        eobrun = 0;
        Bits b{data + pos, data + size};
        auto unit = [&](Component& c, int16_t* blk) -> bool {
            if (!progressive) return decode_sequential(b, c, blk);
            if (dc_scan) {
                if (ah == 0) return decode_dc_first(b, c, blk, al);
                decode_dc_refine(b, blk, al);
                return true;
            }
            return ah == 0 ? decode_ac_first(b, c, blk, ss, se, al) : decode_ac_refine(b, c, blk, ss, se, al);
        };
        int todo = restart_interval;
        auto restart = [&]() -> bool {
            if (!restart_interval) return true;
            if (todo-- > 0) return true;
            if (!b.restart()) return fail(err, "missing restart marker");
            todo = restart_interval - 1;
            for (Component* c : sc) c->dc_pred = 0;
            eobrun = 0;
            return true;
        };
        if (sc.size() == 1) {  // non-interleaved: the component's own blocks, row by row
            Component& c = *sc[0];
            const int nbx = (c.pw + 7) / 8, nby = (c.ph + 7) / 8;
            for (int by = 0; by < nby; ++by)
                for (int bx = 0; bx < nbx; ++bx) {
                    if (!restart()) return false;
                    if (!unit(c, block(c, bx, by))) return false;
                }
        } else {
            for (int my = 0; my < mcuy; ++my)
                for (int mx = 0; mx < mcux; ++mx) {
                    if (!restart()) return false;
                    for (Component* c : sc)
                        for (int v = 0; v < c->v; ++v)
                            for (int h = 0; h < c->h; ++h)
                                if (!unit(*c, block(*c, mx * c->h + h, my * c->v + v))) return false;
                }
        }
        // continue after the entropy-coded data: the next marker that is not RSTn
        const uint8_t* q = b.at_marker ? b.p : data + pos;
        while (q + 1 < data + size && !(q[0] == 0xFF && q[1] != 0x00 && !(q[1] >= 0xD0 && q[1] <= 0xD7) && q[1] != 0xFF))
            ++q;
        pos = (size_t)(q - data);
        return true;
    }

    // dequantise + inverse DCT of every block into the component's sample plane
    void reconstruct(const Component& c, std::vector<uint8_t>& plane) const {
        static double basis[8][8];
        static bool init = false;
        if (!init) {
            for (int u = 0; u < 8; ++u)
                for (int x = 0; x < 8; ++x)
                    basis[x][u] = std::cos((2 * x + 1) * u * M_PI / 16.0) * (u ? 0.5 : 0.5 / std::sqrt(2.0));
            init = true;
        }
        const int stride = c.bw * 8;
        plane.assign((size_t)stride * c.bh * 8, 0);
        for (int by = 0; by < c.bh; ++by)
            for (int bx = 0; bx < c.bw; ++bx) {
                const int16_t* blk = &c.coef[((size_t)by * c.bw + bx) * 64];
                double f[64], tmp[64];
                for (int i = 0; i < 64; ++i) f[i] = (double)blk[i] * c.q[i];
                for (int v = 0; v < 8; ++v)  // rows: inverse over the horizontal frequencies
                    for (int x = 0; x < 8; ++x) {
                        double s = 0;
                        for (int u = 0; u < 8; ++u) s += basis[x][u] * f[v * 8 + u];
                        tmp[v * 8 + x] = s;
                    }
                for (int x = 0; x < 8; ++x)  // columns: inverse over the vertical frequencies
                    for (int y = 0; y < 8; ++y) {
                        double s = 0;
                        for (int v = 0; v < 8; ++v) s += basis[y][v] * tmp[v * 8 + x];
                        const long r = std::lround(s + 128.0);
                        plane[(size_t)(by * 8 + y) * stride + bx * 8 + x] = (uint8_t)std::min(255L, std::max(0L, r));
                    }
            }
    }

    // component plane -> full-resolution samples (width x height)
    void upsample(const Component& c, const std::vector<uint8_t>& plane, std::vector<uint8_t>& out) const {
        const int stride = c.bw * 8, sx = hmax / c.h, sy = vmax / c.v;
        out.assign((size_t)width * height, 0);
        const int cw = c.pw, ch = c.ph;
        auto at = [&](int x, int y) { return (int)plane[(size_t)y * stride + x]; };
        if (sx == 1 && sy == 1) {
            for (int y = 0; y < height; ++y)
                for (int x = 0; x < width; ++x) out[(size_t)y * width + x] = (uint8_t)at(x, y);
        } else if (sx == 2 && sy == 1 && cw > 2) {
            for (int y = 0; y < height; ++y)
                for (int x = 0; x < width; ++x) {
                    const int i = x >> 1;
                    int v;
                    if (x & 1) v = i + 1 < cw ? (3 * at(i, y) + at(i + 1, y) + 2) >> 2 : at(i, y);
                    else v = i > 0 ? (3 * at(i, y) + at(i - 1, y) + 1) >> 2 : at(i, y);
                    out[(size_t)y * width + x] = (uint8_t)v;
                }
        } else if (sx == 2 && sy == 2 && cw > 2) {
            std::vector<int> colsum(cw);
            for (int y = 0; y < height; ++y) {
                const int j = y >> 1;
                // nearer input row weighs 3, the other neighbour 1 (edges repeat)
                const int jn = (y & 1) ? std::min(j + 1, ch - 1) : std::max(j - 1, 0);
                for (int i = 0; i < cw; ++i) colsum[i] = 3 * at(i, j) + at(i, jn);
                for (int x = 0; x < width; ++x) {
                    const int i = x >> 1;
                    int v;
                    if (x & 1) v = i + 1 < cw ? (3 * colsum[i] + colsum[i + 1] + 7) >> 4 : (colsum[i] * 4 + 7) >> 4;
                    else v = i > 0 ? (3 * colsum[i] + colsum[i - 1] + 8) >> 4 : (colsum[i] * 4 + 8) >> 4;
                    out[(size_t)y * width + x] = (uint8_t)v;
                }
            }
        } else {
            for (int y = 0; y < height; ++y)
                for (int x = 0; x < width; ++x) out[(size_t)y * width + x] = (uint8_t)at(x / sx, y / sy);
        }
    }

    bool run(DecodedImage& img) {
        if (size < 4 || data[0] != 0xFF || data[1] != 0xD8) return fail(err, "not a JPEG file");
        size_t pos = 2;
        bool eoi = false;
        int scans = 0;
        while (pos + 1 < size) {
            if (data[pos] != 0xFF) {
                ++pos;  // tolerate padding between segments
                continue;
            }
            const int marker = data[pos + 1];
            pos += 2;
            if (marker == 0xFF || (marker >= 0xD0 && marker <= 0xD7) || marker == 0x01) {
                if (marker == 0xFF) --pos;
                continue;
            }
            if (marker == 0xD9) {
                eoi = true;
                break;
            }
            if (pos + 2 > size) return fail(err, "truncated segment");
            const size_t len = (size_t)(data[pos] << 8 | data[pos + 1]);
            if (len < 2 || pos + len > size) return fail(err, "truncated segment");
            const uint8_t* p = data + pos + 2;
            const size_t n = len - 2;
            pos += len;
            switch (marker) {
                case 0xDB:
                    if (!read_dqt(p, n)) return false;
                    break;
                case 0xC4:
                    if (!read_dht(p, n)) return false;
                    break;
                case 0xC0:
                case 0xC1:
                case 0xC2:
                    if (!read_sof(p, n, marker)) return false;
                    break;
                case 0xC3: case 0xC5: case 0xC6: case 0xC7: case 0xC9: case 0xCA: case 0xCB: case 0xCD: case 0xCE:
                case 0xCF:
                    return fail(err, "arithmetic-coded, lossless and hierarchical JPEG are not supported");
                case 0xDD:
                    if (n < 2) return fail(err, "bad DRI");
                    restart_interval = p[0] << 8 | p[1];
                    break;
                case 0xDA:
                    if (!read_sos(p, n, pos)) return false;
                    ++scans;
                    break;
                case 0xE0:
                    if (n >= 5 && std::memcmp(p, "JFIF\0", 5) == 0) jfif = true;
                    break;
                case 0xEE:
                    if (n >= 12 && std::memcmp(p, "Adobe", 5) == 0) {
                        adobe = true;
                        adobe_transform = p[11];
                    }
                    break;
                default:
                    break;  // APPn, COM and others
            }
        }
        if (!frame || scans == 0) return fail(err, "no image data");
        (void)eoi;  // a missing EOI is tolerated, as decoders commonly do
        std::vector<std::vector<uint8_t>> full(comp.size());
        for (size_t i = 0; i < comp.size(); ++i) {
            if (!comp[i].q_latched) return fail(err, "a component has no scan");
            std::vector<uint8_t> plane;
            reconstruct(comp[i], plane);
            upsample(comp[i], plane, full[i]);
        }
        img.width = (uint32_t)width;
        img.height = (uint32_t)height;
        img.bits = 8;
        const size_t np = (size_t)width * height;
        if (comp.size() == 1) {
            img.channels = 1;
            img.samples.assign(full[0].begin(), full[0].end());
            return true;
        }
        img.channels = 3;
        img.samples.resize(np * 3);
        const bool rgb = (adobe && adobe_transform == 0) ||
                         (!jfif && comp[0].id == 'R' && comp[1].id == 'G' && comp[2].id == 'B');
        for (size_t i = 0; i < np; ++i) {
            if (rgb) {
                for (int k = 0; k < 3; ++k) img.samples[i * 3 + k] = full[k][i];
                continue;
            }
            const double y = full[0][i], cb = full[1][i] - 128.0, cr = full[2][i] - 128.0;
            const double rgbv[3] = {y + 1.402 * cr, y - 0.344136 * cb - 0.714136 * cr, y + 1.772 * cb};
            for (int k = 0; k < 3; ++k) img.samples[i * 3 + k] = (uint16_t)std::min(255L, std::max(0L, std::lround(rgbv[k])));
        }
        return true;
    }
};

}  // namespace

bool decode_jpeg(const uint8_t* data, size_t size, DecodedImage& out, std::string& err) {
    Decoder d(data, size, err);
    return d.run(out);
}

}  // namespace ptcore
