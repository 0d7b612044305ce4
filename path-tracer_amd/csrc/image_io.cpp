// image_io.cpp - decoders for the image files the reference's front end reads:
// PNG (glTF textures and sky images) and Radiance RGBE .hdr (sky images);
// OpenEXR skies are in exr_decode.cpp, JPEG images in jpeg_decode.cpp.
//
// The reference decodes through the `image` crate (Cargo.toml: image 0.25,
// via gltf::import_slice for textures and image::load_from_memory for the
// sky, world.rs:887-889, 1872-1873). That crate is not part of the reference
// tree; these decoders follow the file-format specifications (PNG: ISO/IEC
// 15948, Radiance picture format) and the crate's published conversions:
// 8/16-bit samples to f32 by division by 255/65535, RGBE to f32 as
// c * 2^(e - 136) (e = 0 is black).
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "image_io.h"

namespace ptcore {

namespace {

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

bool fail(std::string& err, const std::string& msg) {
    err = msg;
    return false;
}

int paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

}  // namespace

bool decode_png(const uint8_t* data, size_t size, DecodedImage& out, std::string& err) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (size < 8 || std::memcmp(data, sig, 8) != 0) return fail(err, "not a PNG file");
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, palette, trns;
    bool have_ihdr = false, done = false;
    while (!done) {
        if (pos + 12 > size) return fail(err, "truncated PNG chunk");
        const uint32_t len = be32(data + pos);
        const uint8_t* type = data + pos + 4;
        if (len > size - pos - 12) return fail(err, "truncated PNG chunk");
        const uint8_t* d = data + pos + 8;
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) return fail(err, "bad IHDR");
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
            if (d[10] != 0 || d[11] != 0) return fail(err, "unsupported PNG compression or filter method");
            have_ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            palette.assign(d, d + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(d, d + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            done = true;
        }
        pos += 12 + (size_t)len;
    }
    if (!have_ihdr || w == 0 || h == 0) return fail(err, "PNG without a valid IHDR");
    if (interlace != 0) return fail(err, "interlaced PNG is not supported");
    int channels;
    switch (ctype) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: return fail(err, "bad PNG colour type");
    }
    const bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                          (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                          ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!depth_ok) return fail(err, "bad PNG bit depth");
    if ((uint64_t)w * h > MAX_IMAGE_PIXELS)
        return fail(err, "PNG of " + std::to_string(w) + "x" + std::to_string(h) + " exceeds the image size limit");
    const size_t row_bits = (size_t)w * channels * depth;
    const size_t row_bytes = (row_bits + 7) / 8;
    const size_t bpp = std::max<size_t>(1, (size_t)channels * depth / 8);
    std::vector<uint8_t> raw((row_bytes + 1) * h);
    uLongf raw_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &raw_len, idat.data(), (uLong)idat.size()) != Z_OK || raw_len != raw.size())
        return fail(err, "corrupt PNG image data");
    // unfilter in place into rows
    std::vector<uint8_t> img(row_bytes * h);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t ft = raw[y * (row_bytes + 1)];
        const uint8_t* src = &raw[y * (row_bytes + 1) + 1];
        uint8_t* cur = &img[y * row_bytes];
        const uint8_t* prev = y ? &img[(y - 1) * row_bytes] : nullptr;
        for (size_t i = 0; i < row_bytes; ++i) {
            int a = i >= bpp ? cur[i - bpp] : 0, b = prev ? prev[i] : 0, c = (prev && i >= bpp) ? prev[i - bpp] : 0;
            int x = src[i];
            switch (ft) {
                case 0: break;
                case 1: x += a; break;
                case 2: x += b; break;
                case 3: x += (a + b) >> 1; break;
                case 4: x += paeth(a, b, c); break;
                default: return fail(err, "bad PNG filter type");
            }
            cur[i] = (uint8_t)x;
        }
    }
    out.width = w;
    out.height = h;
    // sample extraction; 16-bit samples are kept whole
    auto sample = [&](uint32_t y, size_t k) -> uint32_t {  // k-th sample of row y
        const uint8_t* r = &img[y * row_bytes];
        if (depth == 16) return (uint32_t)r[2 * k] << 8 | r[2 * k + 1];
        if (depth == 8) return r[k];
        const size_t bit = k * depth;
        return (r[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1u);
    };
    if (ctype == 3) {
        // palette expansion (RGB, or RGBA when tRNS is present), as the decoder crate does
        if (palette.empty() || palette.size() % 3) return fail(err, "PNG palette missing");
        const size_t entries = palette.size() / 3;
        out.channels = trns.empty() ? 3 : 4;
        out.bits = 8;
        out.samples.resize((size_t)w * h * out.channels);
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) {
                uint32_t idx = sample(y, x);
                if (idx >= entries) return fail(err, "PNG palette index out of range");
                uint16_t* o = &out.samples[((size_t)y * w + x) * out.channels];
                for (int c = 0; c < 3; ++c) o[c] = palette[idx * 3 + c];
                if (out.channels == 4) o[3] = idx < trns.size() ? trns[idx] : 255;
            }
        return true;
    }
    out.channels = channels;
    out.bits = depth == 16 ? 16 : 8;
    out.samples.resize((size_t)w * h * channels);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
            for (int c = 0; c < channels; ++c) {
                uint32_t v = sample(y, (size_t)x * channels + c);
                if (depth < 8) v = v * 255u / ((1u << depth) - 1u);  // low-bit grey scaled to 8 bits
                out.samples[((size_t)y * w + x) * channels + c] = (uint16_t)v;
            }
    return true;
}

bool decode_hdr(const uint8_t* data, size_t size, FloatImage& out, std::string& err) {
    size_t pos = 0;
    auto line = [&](std::string& s) {
        s.clear();
        while (pos < size && data[pos] != '\n') s.push_back((char)data[pos++]);
        if (pos >= size) return false;
        ++pos;
        return true;
    };
    std::string s;
    if (!line(s) || (s.rfind("#?", 0) != 0)) return fail(err, "not a Radiance HDR file");
    for (;;) {
        if (!line(s)) return fail(err, "truncated HDR header");
        if (s.empty()) break;
        if (s.rfind("FORMAT=", 0) == 0 && s != "FORMAT=32-bit_rle_rgbe")
            return fail(err, "unsupported HDR pixel format " + s.substr(7));
    }
    if (!line(s)) return fail(err, "missing HDR resolution");
    char ya[3] = {0}, xa[3] = {0};
    long hh = 0, ww = 0;
    if (std::sscanf(s.c_str(), "%2s %ld %2s %ld", ya, &hh, xa, &ww) != 4 || std::strcmp(ya, "-Y") != 0 ||
        std::strcmp(xa, "+X") != 0)
        return fail(err, "unsupported HDR orientation '" + s + "'");
    if (ww <= 0 || hh <= 0 || ww > 65535 || hh > 65535) return fail(err, "bad HDR size");
    if ((uint64_t)ww * (uint64_t)hh > MAX_IMAGE_PIXELS)
        return fail(err, "HDR image of " + std::to_string(ww) + "x" + std::to_string(hh) + " exceeds the image size limit");
    const uint32_t w = (uint32_t)ww, h = (uint32_t)hh;
    out.width = w;
    out.height = h;
    out.rgba.assign((size_t)w * h * 4, 0.0f);
    std::vector<uint8_t> scan((size_t)w * 4);
    for (uint32_t y = 0; y < h; ++y) {
        if (pos + 4 > size) return fail(err, "truncated HDR data");
        const uint8_t* p = data + pos;
        if (w >= 8 && w < 32768 && p[0] == 2 && p[1] == 2 && (((uint32_t)p[2] << 8) | p[3]) == w && !(p[2] & 0x80)) {
            // run-length encoded scanline: each of the 4 components separately
            pos += 4;
            for (int c = 0; c < 4; ++c) {
                uint32_t x = 0;
                while (x < w) {
                    if (pos >= size) return fail(err, "truncated HDR run");
                    uint32_t n = data[pos++];
                    if (n > 128) {
                        n -= 128;
                        if (pos >= size || x + n > w) return fail(err, "bad HDR run");
                        for (uint32_t k = 0; k < n; ++k) scan[(size_t)(x + k) * 4 + c] = data[pos];
                        ++pos;
                    } else {
                        if (n == 0 || x + n > w || pos + n > size) return fail(err, "bad HDR literal run");
                        for (uint32_t k = 0; k < n; ++k) scan[(size_t)(x + k) * 4 + c] = data[pos++];
                    }
                    x += n;
                }
            }
        } else {
            // flat RGBE scanline
            if (pos + (size_t)w * 4 > size) return fail(err, "truncated HDR scanline");
            std::memcpy(scan.data(), data + pos, (size_t)w * 4);
            pos += (size_t)w * 4;
        }
        for (uint32_t x = 0; x < w; ++x) {
            const uint8_t* e = &scan[(size_t)x * 4];
            float* o = &out.rgba[((size_t)y * w + x) * 4];
            if (e[3] != 0) {
                for (int c = 0; c < 3; ++c) o[c] = std::ldexp((float)e[c], (int)e[3] - 136);
            }
            o[3] = 1.0f;
        }
    }
    return true;
}

bool decode_sky_image(const uint8_t* data, size_t size, FloatImage& out, std::string& err) {
    if (size >= 2 && data[0] == '#' && data[1] == '?') return decode_hdr(data, size, out, err);
    const bool png = size >= 8 && data[0] == 137 && data[1] == 'P';
    const bool jpeg = size >= 3 && data[0] == 0xFF && data[1] == 0xD8 && data[2] == 0xFF;
    if (png || jpeg) {
        DecodedImage img;
        if (!(png ? decode_png(data, size, img, err) : decode_jpeg(data, size, img, err))) return false;
        out.width = img.width;
        out.height = img.height;
        out.rgba.assign((size_t)img.width * img.height * 4, 0.0f);
        const float scale = img.bits == 16 ? 65535.0f : 255.0f;
        for (size_t i = 0; i < (size_t)img.width * img.height; ++i) {
            const uint16_t* s = &img.samples[i * img.channels];
            float* o = &out.rgba[i * 4];
            float v[4] = {0.0f, 0.0f, 0.0f, 1.0f};
            switch (img.channels) {  // DynamicImage::into_rgba32f
                case 1: v[0] = v[1] = v[2] = (float)s[0] / scale; break;
                case 2: v[0] = v[1] = v[2] = (float)s[0] / scale; v[3] = (float)s[1] / scale; break;
                case 3: for (int c = 0; c < 3; ++c) v[c] = (float)s[c] / scale; break;
                default: for (int c = 0; c < 4; ++c) v[c] = (float)s[c] / scale; break;
            }
            std::memcpy(o, v, sizeof(v));
        }
        return true;
    }
    if (size >= 4 && data[0] == 0x76 && data[1] == 0x2f && data[2] == 0x31 && data[3] == 0x01)
        return decode_exr(data, size, out, err);
    return fail(err, "unrecognised sky image format (supported: Radiance .hdr, OpenEXR, PNG, JPEG)");
}

}  // namespace ptcore
