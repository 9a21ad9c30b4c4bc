// exr_writer.cpp - OpenEXR output of rendered images as the reference renderer writes them
// (kennyalive/YAR src/ref/reference_renderer.cpp:673-705 write_openexr_image, src/lib/image.cpp:163-224
// Image::write_exr): a single-part scanline file with HALF channels B, G, R, no compression or ZIP
// (16 scanlines per block, zlib), and the yar_* custom attributes (build version, asserts, render
// device, input file, spp, variance; load and render time only when varying attributes are enabled).
// Written from the OpenEXR file-layout specification: magic, version, header attributes
// (name, type, size, value) ending in a null byte, a scanline-block offset table, then the blocks
// (first line y, data size, data).
#include "common.hpp"

#include <zlib.h>

#include <cstdio>
#include <string>

namespace yar {
namespace {

// IEEE binary32 -> binary16, round to nearest even; overflow to infinity, NaN kept quiet.
uint16_t float_to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t absx = x & 0x7fffffffu;
    if (absx >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (absx > 0x7f800000u ? 0x200u | ((absx >> 13) & 0x3ffu) : 0u));
    if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);           // rounds to >= 65520: infinity
    if (absx < 0x38800000u) {                                               // below the smallest normal half
        if (absx < 0x33000000u) return (uint16_t)sign;                      // below half of the smallest subnormal
        const uint32_t m = (absx & 0x7fffffu) | 0x800000u;
        const int e = (int)(absx >> 23);                                    // 102 .. 112
        const int shift = 126 - e;                                          // 14 .. 24
        uint32_t h = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((absx - 0x38000000u) >> 13);                              // rebias exponent 127 -> 15
    const uint32_t rem = absx & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

struct Bytes {
    std::string s;
    void raw(const void* p, size_t n) { s.append(reinterpret_cast<const char*>(p), n); }
    void str0(const std::string& v) { s.append(v); s.push_back('\0'); }
    void i32(int32_t v) { raw(&v, 4); }
    void u8(uint8_t v) { raw(&v, 1); }
    void f32(float v) { raw(&v, 4); }
    void u64(uint64_t v) { raw(&v, 8); }
    void attr(const std::string& name, const std::string& type, const std::string& value) {
        str0(name);
        str0(type);
        i32((int32_t)value.size());
        s.append(value);
    }
};

std::string pod(const void* p, size_t n) { return std::string(reinterpret_cast<const char*>(p), n); }

// OpenEXR ZIP block coding: bytes split into even and odd halves, then a byte-delta predictor, then zlib.
bool zip_block(const std::string& raw, std::string& out) {
    const size_t n = raw.size();
    std::string t(n, '\0');
    size_t a = 0, b = (n + 1) / 2;
    for (size_t i = 0; i < n; i += 2) {
        t[a++] = raw[i];
        if (i + 1 < n) t[b++] = raw[i + 1];
    }
    for (size_t i = n; i-- > 1;) {
        const int d = (int)(uint8_t)t[i] - (int)(uint8_t)t[i - 1] + (128 + 256);
        t[i] = (char)(uint8_t)(d & 0xff);
    }
    uLongf cap = compressBound((uLong)n);
    out.assign(cap, '\0');
    if (compress(reinterpret_cast<Bytef*>(&out[0]), &cap, reinterpret_cast<const Bytef*>(t.data()), (uLong)n) != Z_OK)
        return false;
    out.resize(cap);
    return true;
}

}  // namespace
}  // namespace yar

extern "C" int yar_image_write_exr(const char* path, const float* rgb, int32_t width, int32_t height,
                                   const yar_exr_attributes* attributes, int32_t compress_zip) {
    using namespace yar;
    if (!path || !rgb || width <= 0 || height <= 0) {
        set_last_error("yar_image_write_exr: invalid argument");
        return YAR_ERR_INVALID_ARGUMENT;
    }
    return abi_guard("yar_image_write_exr", [&] {
        yar_exr_attributes at{};
        if (attributes) at = *attributes;
        Bytes h;
        const uint32_t magic = 20000630u;                                  // 0x762f3101
        h.raw(&magic, 4);
        h.i32(2);                                                          // version 2, single-part scanline
        {   // channels: B, G, R (alphabetical), HALF, not linear, sampling 1
            Bytes c;
            for (const char* name : {"B", "G", "R"}) {
                c.str0(name);
                c.i32(1);
                c.u8(0); c.u8(0); c.u8(0); c.u8(0);
                c.i32(1); c.i32(1);
            }
            c.u8(0);
            h.attr("channels", "chlist", c.s);
        }
        const uint8_t compression = compress_zip ? 3 : 0;                  // ZIP_COMPRESSION / NO_COMPRESSION
        h.attr("compression", "compression", pod(&compression, 1));
        const int32_t box[4] = {0, 0, width - 1, height - 1};
        h.attr("dataWindow", "box2i", pod(box, 16));
        h.attr("displayWindow", "box2i", pod(box, 16));
        const uint8_t line_order = 0;                                      // INCREASING_Y
        h.attr("lineOrder", "lineOrder", pod(&line_order, 1));
        const float aspect = 1.f, center[2] = {0.f, 0.f}, sww = 1.f;
        h.attr("pixelAspectRatio", "float", pod(&aspect, 4));
        h.attr("screenWindowCenter", "v2f", pod(center, 8));
        h.attr("screenWindowWidth", "float", pod(&sww, 4));
        // custom attributes of write_openexr_image, in its order
        std::string dump;
        auto add_s = [&](const char* name, const std::string& v, bool to_image) {
            if (to_image) h.attr(name, "string", v);
            dump += std::string(name) + " " + v + "\n";
        };
        auto add_i = [&](const char* name, int32_t v, bool to_image) {
            if (to_image) h.attr(name, "int", pod(&v, 4));
            dump += std::string(name) + " " + std::to_string(v) + "\n";
        };
        auto add_f = [&](const char* name, float v, bool to_image) {
            if (to_image) h.attr(name, "float", pod(&v, 4));
            char buf[64];
            std::snprintf(buf, sizeof(buf), "%f", v);
            dump += std::string(name) + " " + buf + "\n";
        };
        add_s("yar_build_version", "0.0", true);
        add_i("yar_build_asserts", 0, true);
        add_s("yar_render_device", "gpu", true);
        add_s("yar_input_file", at.input_file ? at.input_file : "", true);
        add_i("yar_spp", at.spp, true);
        add_f("yar_variance", at.variance, true);
        add_f("yar_load_time", at.load_time, at.enable_varying_attributes != 0);
        add_f("yar_render_time", at.render_time, at.enable_varying_attributes != 0);
        h.u8(0);                                                           // end of header

        const int lines_per_block = compress_zip ? 16 : 1;
        const int n_blocks = (height + lines_per_block - 1) / lines_per_block;
        std::vector<std::string> blocks((size_t)n_blocks);
        for (int bi = 0; bi < n_blocks; bi++) {
            const int y0 = bi * lines_per_block, y1 = std::min(height, y0 + lines_per_block);
            std::string raw;
            raw.reserve((size_t)(y1 - y0) * width * 6);
            for (int y = y0; y < y1; y++)
                for (int ch = 2; ch >= 0; ch--)                            // B, G, R
                    for (int x = 0; x < width; x++) {
                        const uint16_t v = float_to_half(rgb[3 * ((size_t)y * width + x) + ch]);
                        raw.append(reinterpret_cast<const char*>(&v), 2);
                    }
            std::string data = raw;
            if (compress_zip) {
                std::string z;
                if (!zip_block(raw, z)) throw Error{YAR_ERR_INTERNAL, "zlib compression failed"};
                if (z.size() < raw.size()) data.swap(z);                   // stored raw when zip does not shrink it
            }
            Bytes blk;
            blk.i32(y0);
            blk.i32((int32_t)data.size());
            blk.s += data;
            blocks[(size_t)bi] = std::move(blk.s);
        }
        uint64_t off = h.s.size() + 8ull * (uint64_t)n_blocks;
        for (const std::string& b : blocks) {
            h.u64(off);
            off += b.size();
        }
        FILE* f = std::fopen(path, "wb");
        if (!f) throw Error{YAR_ERR_IO, std::string("cannot write ") + path};
        bool ok = std::fwrite(h.s.data(), 1, h.s.size(), f) == h.s.size();
        for (const std::string& b : blocks) ok = ok && std::fwrite(b.data(), 1, b.size(), f) == b.size();
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) throw Error{YAR_ERR_IO, std::string("write failed: ") + path};
        if (at.dump_attributes) {                                          // <image>.txt beside the image
            std::string p(path);
            const size_t dot = p.find_last_of('.'), slash = p.find_last_of('/');
            if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) p.resize(dot);
            p += ".txt";
            FILE* d = std::fopen(p.c_str(), "w");
            if (!d) throw Error{YAR_ERR_IO, "cannot write " + p};
            const bool dok = std::fwrite(dump.data(), 1, dump.size(), d) == dump.size();
            if (std::fclose(d) != 0 || !dok) throw Error{YAR_ERR_IO, "write failed: " + p};
        }
    });
}
