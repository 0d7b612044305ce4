// image_io.h - image decoders of the asset front end (PNG and JPEG textures, sky images).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace ptcore {

// Largest image any decoder accepts (8192 x 8192 pixels), checked against the
// header before anything is allocated: a small file cannot make the loader
// allocate or process an arbitrarily large image. (The reference's `image`
// crate refuses decodes beyond its 512 MiB default allocation limit.)
constexpr uint64_t MAX_IMAGE_PIXELS = 1ull << 26;

// Decoded integer image: `channels` interleaved samples per pixel, 8 or 16
// significant bits each (palettes expanded to RGB/RGBA, low-bit grey scaled to 8).
struct DecodedImage {
    uint32_t width = 0, height = 0;
    int channels = 0, bits = 8;
    std::vector<uint16_t> samples;
};

// RGBA32F image, rows top to bottom.
struct FloatImage {
    uint32_t width = 0, height = 0;
    std::vector<float> rgba;
};

bool decode_png(const uint8_t* data, size_t size, DecodedImage& out, std::string& err);
// Baseline/extended/progressive Huffman JPEG, 1 or 3 components, to 8-bit grey or RGB.
bool decode_jpeg(const uint8_t* data, size_t size, DecodedImage& out, std::string& err);
bool decode_hdr(const uint8_t* data, size_t size, FloatImage& out, std::string& err);
// Single-part scanline OpenEXR (NONE, ZIPS, ZIP, PIZ; HALF/FLOAT/UINT channels) to RGBA32F.
bool decode_exr(const uint8_t* data, size_t size, FloatImage& out, std::string& err);
// A sky image as image::load_from_memory(..).into_rgba32f() gives it (Radiance .hdr, OpenEXR, PNG or JPEG).
bool decode_sky_image(const uint8_t* data, size_t size, FloatImage& out, std::string& err);

// PNG of 8-bit rows with 3 (RGB) or 4 (RGBA) channels.
bool encode_png_channels(const uint8_t* pixels, size_t stride, uint32_t w, uint32_t h, int channels,
                         std::vector<uint8_t>& out, std::string& err);
// Uncompressed scanline OpenEXR with FLOAT R, G, B, A channels (lossless for RGBA32F).
bool encode_exr_float(const float* rgba, uint32_t w, uint32_t h, std::vector<uint8_t>& out, std::string& err);
// Radiance .hdr (run-length coded RGBE scanlines) of an image whose every value
// the RGBE decoding c * 2^(E - 136) gives back exactly and whose alpha is 1;
// fails otherwise.
bool encode_hdr_exact(const float* rgba, uint32_t w, uint32_t h, std::vector<uint8_t>& out, std::string& err);

}  // namespace ptcore
