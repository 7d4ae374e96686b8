// OpenEXR output of the accumulated image, the reference's saveFrame
// (host/Renderer.swift:804-807 -> utils/Metal.swift:75-124, which calls
// tinyexr's SaveEXR with four FLOAT channels).
//
// What is written: RGB times `norm` (float multiply, as the reference scales
// its buffer), alpha 1, as 32-bit float channels A, B, G, R (the channel list
// of an EXR file is sorted by name) in one part of scanline blocks, one
// scanline per block, increasing y, no compression.  The reference's tinyexr
// compresses larger images with ZIP; the pixel values a reader decodes are
// the same either way.  The layout follows the OpenEXR file format document
// (magic number, version 2, attribute header, line offset table, blocks of
// y, byte count and per-channel scanline data, all little-endian).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/rm_host.h"

namespace {

struct Bytes {
    std::vector<uint8_t> b;
    void u8(uint8_t v) { b.push_back(v); }
    void i32(int32_t v) {
        uint32_t u;
        memcpy(&u, &v, 4);
        for (int k = 0; k < 4; k++) b.push_back((uint8_t)(u >> (8 * k)));
    }
    void u64(uint64_t v) {
        for (int k = 0; k < 8; k++) b.push_back((uint8_t)(v >> (8 * k)));
    }
    void f32(float v) {
        uint32_t u;
        memcpy(&u, &v, 4);
        for (int k = 0; k < 4; k++) b.push_back((uint8_t)(u >> (8 * k)));
    }
    void str(const char* s) {
        while (*s) b.push_back((uint8_t)*s++);
        b.push_back(0);
    }
    // attribute: name, type name, byte size, value (written by `value`)
    template <class F>
    void attr(const char* name, const char* type, int32_t size, F value) {
        str(name);
        str(type);
        i32(size);
        const size_t at = b.size();
        value();
        if (b.size() - at != (size_t)size) throw std::logic_error(std::string("attribute size mismatch: ") + name);
    }
};

void write_exr(const char* path, const float* rgba, uint32_t width, uint32_t height, float norm) {
    if (width == 0 || height == 0) throw std::runtime_error("empty image");
    // a block's byte count (16 bytes per pixel of a scanline) and the window
    // corners are 32-bit signed integers in the file
    if (16ull * width > 0x7fffffffull || height > 0x7fffffffu) throw std::runtime_error("image too large");
    static const char* kChannels[4] = {"A", "B", "G", "R"};  // sorted by name
    static const int kSource[4] = {3, 2, 1, 0};              // rgba component of each channel

    Bytes h;
    h.i32(20000630);  // magic
    h.i32(2);         // version 2, single-part scanline, short names
    h.attr("channels", "chlist", 4 * (2 + 16) + 1, [&] {
        for (const char* c : kChannels) {
            h.str(c);
            h.i32(2);  // FLOAT
            h.u8(0);   // pLinear
            h.u8(0);
            h.u8(0);
            h.u8(0);  // reserved
            h.i32(1);  // x sampling
            h.i32(1);  // y sampling
        }
        h.u8(0);
    });
    h.attr("compression", "compression", 1, [&] { h.u8(0); });  // NO_COMPRESSION
    h.attr("dataWindow", "box2i", 16, [&] {
        h.i32(0);
        h.i32(0);
        h.i32((int32_t)width - 1);
        h.i32((int32_t)height - 1);
    });
    h.attr("displayWindow", "box2i", 16, [&] {
        h.i32(0);
        h.i32(0);
        h.i32((int32_t)width - 1);
        h.i32((int32_t)height - 1);
    });
    h.attr("lineOrder", "lineOrder", 1, [&] { h.u8(0); });  // INCREASING_Y
    h.attr("pixelAspectRatio", "float", 4, [&] { h.f32(1.0f); });
    h.attr("screenWindowCenter", "v2f", 8, [&] {
        h.f32(0.0f);
        h.f32(0.0f);
    });
    h.attr("screenWindowWidth", "float", 4, [&] { h.f32(1.0f); });
    h.u8(0);  // end of header

    const uint64_t line_bytes = 4ull * 4ull * width;
    const uint64_t block_bytes = 8 + line_bytes;
    const uint64_t first = h.b.size() + 8ull * height;
    for (uint32_t y = 0; y < height; y++) h.u64(first + (uint64_t)y * block_bytes);

    FILE* f = fopen(path, "wb");
    if (!f) throw std::runtime_error(std::string("could not open '") + path + "' for writing");
    bool ok = fwrite(h.b.data(), 1, h.b.size(), f) == h.b.size();
    Bytes line;
    for (uint32_t y = 0; ok && y < height; y++) {
        line.b.clear();
        line.i32((int32_t)y);
        line.i32((int32_t)line_bytes);
        const float* row = rgba + 4ull * width * y;
        for (int c = 0; c < 4; c++)
            for (uint32_t x = 0; x < width; x++)
                line.f32(kSource[c] == 3 ? 1.0f : row[4ull * x + kSource[c]] * norm);
        ok = fwrite(line.b.data(), 1, line.b.size(), f) == line.b.size();
    }
    if (fclose(f) != 0) ok = false;
    if (!ok) throw std::runtime_error(std::string("write to '") + path + "' failed");
}

}  // namespace

extern "C" int rmh_save_exr(const char* path, const float* rgba, uint32_t width, uint32_t height, float norm,
                            char** out_error) {
    if (out_error) *out_error = nullptr;
    if (!path || !rgba) return -1;
    try {
        write_exr(path, rgba, width, height, norm);
        return 0;
    } catch (const std::exception& e) {
        if (out_error) *out_error = strdup(e.what());
        return -1;
    }
}
